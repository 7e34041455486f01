#!/usr/bin/env python3
"""Statistics of a fault-injection campaign from its ensemble.csv, as the
reference's scripts/ensemble/aggregate.py computes them. Nothing is re-run.

    python -m mnhack8_amd.aggregate --csv DIR/ensemble.csv --out DIR/agg

Writes into --out (same files, columns and number formats as the reference,
scripts/ensemble/aggregate.py:37-95):
  sensitivity.csv  per (level, bit class): P(benign / transient / persistent /
                   detected), P(not benign) with its Wilson 95 % interval,
                   median and 90th percentile of the finite E_fault_max
  detection.csv    per (level, bit class), over the non-benign runs with
                   detectors on: the fraction caught by D1, D2, D3, any of
                   them (or classed detected), and with more CG iterations
  contrast.csv     per sigma: runs, mean k_max / k_min, mean CG iterations of
                   the clean run, P(not benign) with its interval, median E_max
and, when matplotlib is importable, sensitivity.png, healing.png and
contrast.png. Only rows with status "ok" count.
"""
from __future__ import annotations

import argparse
import csv
import math
import os
import sys
from collections import defaultdict

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from mnhack8_amd.validations import fnum, read_csv  # noqa: E402

LEVELS = {"1": "G", "2": "R", "3": "H"}
LEVEL_ORDER = ["G", "R", "H"]
CLASSES = ["mant_lo", "mant_hi", "exp", "sign"]
OUTCOMES = ["benign", "transient", "persistent", "detected"]


def wilson(k, n, z=1.96):
    """Proportion k/n with its Wilson score interval, clipped to [0, 1]
    (scripts/ensemble/aggregate.py:24-28); NaNs for n = 0."""
    if n == 0:
        return (float("nan"), float("nan"), float("nan"))
    p = k / n
    den = 1 + z * z / n
    centre = (p + z * z / (2 * n)) / den
    half = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / den
    return (p, max(0.0, centre - half), min(1.0, centre + half))


def _finite_emax(rs):
    e = np.array([fnum(r["E_fault_max"]) for r in rs])
    return e[np.isfinite(e)]


def _fired(r, col):
    return int(r[col]) >= 0


def sensitivity(rows):
    """{(level, class): [rows]} and the sensitivity.csv table."""
    cells = defaultdict(list)
    for r in rows:
        cells[(LEVELS[r["fi_level"]], r["bit_class"])].append(r)
    table, P = [], {}
    for lv in LEVEL_ORDER:
        for bc in CLASSES:
            rs = cells.get((lv, bc), [])
            n = len(rs)
            if not n:
                continue
            cnt = {c: sum(r["class"] == c for r in rs) for c in OUTCOMES}
            p, lo, hi = wilson(n - cnt["benign"], n)
            em = _finite_emax(rs)
            table.append([lv, bc, n] + [f"{cnt[c] / n:.3f}" for c in OUTCOMES] +
                         [f"{p:.3f}", f"{lo:.3f}", f"{hi:.3f}",
                          f"{np.median(em):.3e}" if em.size else "",
                          f"{np.percentile(em, 90):.3e}" if em.size else ""])
            P[(lv, bc)] = (p, lo, hi, n)
    return cells, table, P


def detection(cells):
    table = []
    for lv in LEVEL_ORDER:
        for bc in CLASSES:
            rs = [r for r in cells.get((lv, bc), []) if r["class"] != "benign" and r.get("detect") == "1"]
            n = len(rs)
            if not n:
                continue
            d1 = sum(_fired(r, "d1_first") for r in rs)
            d2 = sum(_fired(r, "d2_first") for r in rs)
            d3 = sum(_fired(r, "d3_first") for r in rs)
            anyd = sum(_fired(r, "d1_first") or _fired(r, "d2_first") or _fired(r, "d3_first")
                       or r["class"] == "detected" for r in rs)
            cg = sum(int(r["iters_delta_max"]) > 0 for r in rs)
            table.append([lv, bc, n] + [f"{x / n:.2f}" for x in (d1, d2, d3, anyd, cg)])
    return table


def contrast(rows):
    by_sigma = defaultdict(list)
    for r in rows:
        by_sigma[float(r["sigma"])].append(r)
    table = []
    for s in sorted(by_sigma):
        rs = by_sigma[s]
        n = len(rs)
        p, lo, hi = wilson(sum(r["class"] != "benign" for r in rs), n)
        em = _finite_emax(rs)
        ratio = np.mean([fnum(r["kmax"]) / fnum(r["kmin"]) for r in rs])
        cg = np.mean([fnum(r["mean_cg_iters_A"]) for r in rs])
        table.append([s, n, f"{ratio:.3g}", f"{cg:.1f}", f"{p:.3f}", f"{lo:.3f}", f"{hi:.3f}",
                      f"{np.median(em):.3e}" if em.size else ""])
    return by_sigma, table


def _write(path, header, table):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(table)


def _figures(out, rows, P, by_sigma):
    """Three figures under the reference's file names: per-level bars of
    P(not benign) by bit class with their Wilson intervals; the cumulative
    distribution of the healing ratio E_final / E_max per outcome class; and
    P(not benign) and the clean CG work against the field's contrast."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    levels = [lv for lv in LEVEL_ORDER if any(k[0] == lv for k in P)]
    fig, ax = plt.subplots(figsize=(7, 3.2))
    width = 0.8 / max(1, len(levels))
    for j, lv in enumerate(levels):
        xs, ps, err = [], [], [[], []]
        for i, bc in enumerate(CLASSES):
            if (lv, bc) in P:
                p, lo, hi, _ = P[(lv, bc)]
                xs.append(i + (j - (len(levels) - 1) / 2) * width)
                ps.append(p)
                err[0].append(p - lo)
                err[1].append(hi - p)
        ax.bar(xs, ps, width, yerr=err, capsize=2, label=f"level {lv}")
    ax.set_xticks(range(len(CLASSES)), CLASSES)
    ax.set_ylim(0, 1.25)
    ax.set_ylabel("P(not benign), 95% interval")
    ax.legend(fontsize=8, ncol=3, loc="upper center")
    fig.tight_layout()
    fig.savefig(os.path.join(out, "sensitivity.png"), dpi=120)
    plt.close(fig)

    fig, ax = plt.subplots(figsize=(5, 3.2))
    for c in ("benign", "transient", "persistent"):
        ratio = sorted(np.log10(max(fnum(r["E_fault_final"]), 1e-300) / fnum(r["E_fault_max"]))
                       for r in rows if r["class"] == c and fnum(r["E_fault_max"]) > 0)
        if ratio:
            ax.step(ratio, np.arange(1, len(ratio) + 1) / len(ratio), where="post", label=f"{c} ({len(ratio)})")
    ax.set_xlabel("log10(E_final / E_max)")
    ax.set_ylabel("fraction of runs")
    ax.legend(fontsize=8)
    fig.tight_layout()
    fig.savefig(os.path.join(out, "healing.png"), dpi=120)
    plt.close(fig)

    sig = sorted(by_sigma)
    contrast = [np.mean([fnum(r["kmax"]) / fnum(r["kmin"]) for r in by_sigma[s]]) for s in sig]
    stats = [wilson(sum(r["class"] != "benign" for r in by_sigma[s]), len(by_sigma[s])) for s in sig]
    cg = [np.mean([fnum(r["mean_cg_iters_A"]) for r in by_sigma[s]]) for s in sig]
    fig, (top, bottom) = plt.subplots(2, 1, sharex=True, figsize=(4.5, 4.2))
    top.fill_between(contrast, [lo for _, lo, _ in stats], [hi for _, _, hi in stats], alpha=0.3)
    top.plot(contrast, [p for p, _, _ in stats], "o-")
    top.set_ylabel("P(not benign)")
    bottom.plot(contrast, cg, "s-")
    bottom.set_ylabel("CG iterations per step")
    bottom.set_xscale("log")
    bottom.set_xlabel("kmax / kmin")
    fig.tight_layout()
    fig.savefig(os.path.join(out, "contrast.png"), dpi=120)
    plt.close(fig)


def aggregate(csv_path, out, figures=True):
    os.makedirs(out, exist_ok=True)
    rows = [r for r in read_csv(csv_path) if r.get("status") == "ok"]
    print(f"{len(rows)} successful runs")
    cells, sens, P = sensitivity(rows)
    _write(os.path.join(out, "sensitivity.csv"),
           ["level", "bit_class", "n", "p_benign", "p_transient", "p_persistent", "p_detected", "p_not_benign",
            "ci_lo", "ci_hi", "E_max_median", "E_max_p90"], sens)
    _write(os.path.join(out, "detection.csv"),
           ["level", "bit_class", "n_not_benign", "frac_D1", "frac_D2", "frac_D3", "frac_any", "frac_cg_iters"],
           detection(cells))
    by_sigma, con = contrast(rows)
    _write(os.path.join(out, "contrast.csv"),
           ["sigma", "n", "contrast_kmax_kmin", "mean_cg_iters", "p_not_benign", "ci_lo", "ci_hi", "E_max_median"],
           con)
    if figures:
        try:
            _figures(out, rows, P, by_sigma)
        except Exception as ex:  # noqa: BLE001 -- figures are optional, as in the reference
            print("plot skipped:", ex)
    print("wrote", sorted(os.listdir(out)))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--csv", required=True)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    aggregate(a.csv, a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
