#!/usr/bin/env python3
"""Copy the reference's own recorded validation outputs (data files: CSV,
JSON, text) used as golden vectors into tests/golden/v10/ and
tests/golden/ensemble/.

Sources: tests/baseline_validation/v10 of mariagarciapenalva/mnhack8 (the
all-PASS run recorded on an RTX 3060 with CUDA 13.0) and its two recorded
fault-injection campaigns, results/ensemble_N32 and results/ensemble_medium
(ensemble.csv, trimmed to the columns used here, and agg/*.csv).  Only output
data is copied; no source file of the reference is.

    python tests/golden/collect_fixtures.py /root/reference
"""
import os
import shutil
import sys

FILES = [
    "ln16_meta.json", "v6/ln32_meta.json",
    "v0_fast_colored_np1/verify.csv", "v0_fast_atomic_np1/verify.csv",
    "v0_gauss_colored_np1/verify.csv", "v0_gauss_atomic_np1/verify.csv",
    "v1/report.json", "v2/report.json",
    "v2/colored_00/fi_summary_nf.csv", "v2/colored_00/fi_trace_nf.csv",
    "v3/report.json", "v4/report.json", "v4/verify_np1/verify.csv",
    "v5/report.json", "v5/fast/fi_summary_eq.csv", "v5/fast/fi_trace_eq.csv",
    "v5/gauss/fi_summary_eq.csv", "v5/gauss/fi_trace_eq.csv",
    "v6/report.json", "v6/coverage.csv",
    "v6/clean_colored/fi_summary_det.csv", "v6/clean_colored/fi_trace_det.csv",
    "v6/clean_atomic/fi_summary_det.csv", "v6/clean_atomic/fi_trace_det.csv",
]
for lv in (1, 2):
    for b in (30, 51, 55, 62, 63):
        FILES += [f"v6/L{lv}_b{b}/fi_summary_det.csv", f"v6/L{lv}_b{b}/fi_trace_det.csv"]
    for b in (0, 30, 51, 55, 62, 63):
        FILES += [f"v3/L{lv}_b{b}/fi_summary_fi.csv"]
FILES += ["v3/G_boundary/fi_summary_fi.csv", "v3/G_boundary_reproject/fi_summary_fi.csv"]


# The two recorded fault-injection campaigns (results/<name>/ensemble.csv of
# scripts/ensemble/run_ensemble.py, N=32, with agg/ of scripts/ensemble/aggregate.py).
# Only the columns that the grid, the statistics and the GPU comparison read are kept.
CAMPAIGNS = ["ensemble_N32", "ensemble_medium"]
CAMPAIGN_COLUMNS = ["run_id", "field", "sigma", "seed", "bit_class", "status", "fi_level", "fi_bit", "fi_step",
                    "fi_target", "class", "E_fault_max", "E_fault_final", "detect", "d1_first", "d2_first",
                    "d3_first", "iters_delta_max", "kmin", "kmax", "mean_cg_iters_A", "step_at_max",
                    "detected_by", "target_on_boundary"]


def collect_campaigns(ref_root, here):
    import csv
    for name in CAMPAIGNS:
        src = os.path.join(ref_root, "results", name)
        dst = os.path.join(here, "ensemble", name)
        os.makedirs(os.path.join(dst, "agg"), exist_ok=True)
        with open(os.path.join(src, "ensemble.csv"), newline="") as f:
            rows = list(csv.DictReader(f))
        with open(os.path.join(dst, "ensemble.csv"), "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=CAMPAIGN_COLUMNS, extrasaction="ignore", lineterminator="\n")
            w.writeheader()
            w.writerows(rows)
        for fn in ("sensitivity.csv", "detection.csv", "contrast.csv"):
            shutil.copyfile(os.path.join(src, "agg", fn), os.path.join(dst, "agg", fn))
        print(f"{name}: {len(rows)} rows -> {dst}")


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    src = os.path.join(ref_root, "tests", "baseline_validation", "v10")
    dst = os.path.join(here, "v10")
    for rel in FILES:
        os.makedirs(os.path.dirname(os.path.join(dst, rel)), exist_ok=True)
        shutil.copyfile(os.path.join(src, rel), os.path.join(dst, rel))
    print(f"copied {len(FILES)} files to {dst}")
    collect_campaigns(ref_root, here)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
