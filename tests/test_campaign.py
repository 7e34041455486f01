"""The fault-injection campaign (mnhack8_amd/campaign.py, mnhack8_amd/aggregate.py)
against the reference's two recorded campaigns, and its runner.

Golden data: tests/golden/ensemble/ensemble_{N32,medium} (results/<name>/ of the
reference: ensemble.csv trimmed to the columns used here, and agg/*.csv),
copied by tests/golden/collect_fixtures.py.
* The grid: run ids, targets, bits and steps in the recorded order, from the
  same seeded draws (exact).
* The statistics: the three agg tables recomputed from the recorded
  ensemble.csv, character for character (line ends aside).
* The runner, on the CPU: one process per run through the oracle's command
  line (the reference's protocol), resumption, shards and error rows; the
  driver's --campaign file format. The GPU tests compare the campaign's
  driver processes (--campaign) with this protocol.
"""
import csv
import os
import subprocess

import pytest

from mnhack8_amd import aggregate, build, campaign
from oracle_lib import ORACLE_CLI

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ensemble")
# the settings that the recorded campaigns ran with (their run ids and rows
# give N, steps, levels, reps and targets; rng and t/dt are the defaults)
RECORDED = {"ensemble_N32": dict(seeds=[1], reps=2, targets=2),
            "ensemble_medium": dict(seeds=[1, 2], reps=2, targets=3)}


def gold_rows(name):
    with open(os.path.join(GOLD, name, "ensemble.csv"), newline="") as f:
        return list(csv.DictReader(f))


@pytest.mark.parametrize("name", sorted(RECORDED))
def test_grid_reproduces_recorded_campaign(name):
    s = RECORDED[name]
    grid = campaign.build_grid(32, 0.02, 5e-4, [0.5, 1.0, 2.0], s["seeds"], s["reps"], s["targets"], [1, 2, 3],
                               2, 12345)
    rows = gold_rows(name)
    assert [campaign.run_id(g) for g in grid] == [r["run_id"] for r in rows]
    for g, r in zip(grid, rows):
        assert (str(g["level"]), str(g["bit"]), str(g["step"]), str(g["target"])) == \
               (r["fi_level"], r["fi_bit"], r["fi_step"], r["fi_target"])
        assert (g["field"], g["bit_class"], str(g["sigma"]), str(g["seed"])) == \
               (r["field"], r["bit_class"], r["sigma"], r["seed"])


def test_grid_shape_and_draw_ranges():
    grid = campaign.build_grid(16, 0.02, 5e-4, [1.0], [1, 2], 3, 3, [1, 2, 3], 1, 7)
    # mpi 1 drops level H: 2 fields x 2 levels x 4 classes x 3 reps x 3 steps x 3 targets
    assert len(grid) == 2 * 2 * 4 * 3 * 3 * 3
    assert campaign.injection_steps(0.02, 5e-4) == [4, 20, 36]
    assert campaign.injection_steps(0.001, 5e-4) == [1]
    for g in grid:
        assert g["bit"] in campaign.CLASSES[g["bit_class"]]
        n1 = 17
        if g["level"] == 1:
            ix, iy, iz = g["target"] // (n1 * n1), (g["target"] // n1) % n1, g["target"] % n1
            assert all(2 <= c <= 14 for c in (ix, iy, iz))
        else:
            ix, iy, iz = g["target"] // 256, (g["target"] // 16) % 16, g["target"] % 16
            assert all(2 <= c <= 13 for c in (ix, iy, iz))


@pytest.mark.parametrize("name", sorted(RECORDED))
def test_aggregate_reproduces_recorded_statistics(tmp_path, name):
    aggregate.aggregate(os.path.join(GOLD, name, "ensemble.csv"), str(tmp_path), figures=False)
    for fn in ("sensitivity.csv", "detection.csv", "contrast.csv"):
        mine = open(tmp_path / fn, newline="").read().replace("\r\n", "\n")
        gold = open(os.path.join(GOLD, name, "agg", fn), newline="").read().replace("\r\n", "\n")
        assert mine == gold, fn


def test_wilson_interval_properties():
    """Closed forms: k = 0 and k = n clip to 0 and 1, the interval mirrors
    under k -> n - k, and one recorded cell (10 of 36)."""
    p, lo, hi = aggregate.wilson(0, 20)
    assert p == 0 and lo == 0.0 and 0 < hi < 0.2
    p, lo, hi = aggregate.wilson(20, 20)
    assert p == 1 and hi == 1.0 and 0.8 < lo < 1
    for k in range(0, 37):
        a, b = aggregate.wilson(k, 36), aggregate.wilson(36 - k, 36)
        assert abs(a[1] - (1 - b[2])) < 1e-15 and abs(a[2] - (1 - b[1])) < 1e-15
    p, lo, hi = aggregate.wilson(10, 36)
    assert (f"{p:.3f}", f"{lo:.3f}", f"{hi:.3f}") == ("0.278", "0.158", "0.440")
    assert all(x != x for x in aggregate.wilson(0, 0))


def test_ensemble_csv_resume_and_error_rows(tmp_path):
    path = str(tmp_path / "ensemble.csv")
    out = campaign.EnsembleCsv(path)
    assert out.done() == set()
    out.append({"run_id": "a", "field": "f", "sigma": 1.0, "seed": 1, "bit_class": "exp", "status": "ok",
                "class": "benign", "fi_level": "1"})
    g = dict(field="f", sigma=1.0, seed=1, level=3, bit_class="sign", bit=63, step=4, tidx=0)
    out.append(campaign.error_row(g, "no"))
    again = campaign.EnsembleCsv(path)
    assert again.done() == {"a", campaign.run_id(g)}
    rows = list(csv.DictReader(open(path)))
    assert rows[1]["status"] == "error: no" and rows[1]["class"] == "" and rows[1]["fi_level"] == "3"


@pytest.fixture(scope="module")
def small_campaign(tmp_path_factory):
    """A 12^3 campaign (one field, levels G, R, H) run one process per run
    through the oracle's command line."""
    out = str(tmp_path_factory.mktemp("camp"))
    args = ["--out", out, "--N", "12", "--t", "0.005", "--dt", "5e-4", "--sigmas", "1.0", "--seeds", "3",
            "--reps", "1", "--targets", "1", "--detect", "--per-run", "--bin", ORACLE_CLI]
    return out, args


def test_campaign_per_run_protocol(small_campaign):
    out, args = small_campaign
    assert campaign.main(args + ["--shard", "0/2"]) == 0
    rows = list(csv.DictReader(open(os.path.join(out, "ensemble.csv"))))
    grid = campaign.build_grid(12, 0.005, 5e-4, [1.0], [3], 1, 1, [1, 2, 3], 2, 12345)
    assert [r["run_id"] for r in rows] == [campaign.run_id(g) for g in grid[0::2]]
    for r in rows:
        if r["fi_level"] == "3":
            assert r["status"].startswith("error: command failed (rc=1)")
        else:
            assert r["status"] == "ok" and r["class"] in ("benign", "transient", "persistent", "detected")
            d = os.path.join(out, "runs", r["run_id"])
            assert sorted(os.listdir(d)) == ["fi_summary_e.csv", "fi_trace_e.csv", "meta_e.txt"]
    # the other shard, then a rerun that finds nothing left to do
    assert campaign.main(args + ["--shard", "1/2"]) == 0
    assert campaign.main(args) == 0
    rows = list(csv.DictReader(open(os.path.join(out, "ensemble.csv"))))
    assert sorted(r["run_id"] for r in rows) == sorted(campaign.run_id(g) for g in grid)


def test_driver_campaign_flag_checks(tmp_path):
    base = [build.DRIVER, "8", "0.01", "1e-3", str(tmp_path / "f"), str(tmp_path), "--campaign", str(tmp_path / "c")]
    r = subprocess.run(base + ["--fi", "1", "1", "0", "5"], capture_output=True, text=True)
    assert r.returncode == 1 and "--campaign takes its faults from FILE" in r.stderr
    r = subprocess.run(base + ["--dump-snaps"], capture_output=True, text=True)
    assert r.returncode == 1 and "--campaign takes its faults from FILE" in r.stderr
    r = subprocess.run([build.DRIVER], capture_output=True, text=True)
    assert "--campaign FILE" in r.stderr
    # a run directory may contain spaces: DIR is all before the last four fields
    spec = tmp_path / "spec.txt"
    spec.write_text(f"{tmp_path}/run one 1 4 123 55\n{tmp_path}/two 2 20 77 63\n\n")
    r = subprocess.run(base[:6] + ["--campaign", str(spec)], capture_output=True, text=True)
    assert f"Campaign: 2 fault runs from {spec} (first in {tmp_path}/run one)" in r.stdout
    spec.write_text(f"{tmp_path}/x 1 4 123\n")
    r = subprocess.run(base[:6] + ["--campaign", str(spec)], capture_output=True, text=True)
    assert r.returncode == 1 and "every line must read DIR LEVEL STEP TARGET BIT" in r.stderr


def test_aggregate_figures(tmp_path):
    """The three figures of the reference's aggregate step, when matplotlib is there."""
    pytest.importorskip("matplotlib")
    aggregate.aggregate(os.path.join(GOLD, "ensemble_N32", "ensemble.csv"), str(tmp_path), figures=True)
    for fn in ("sensitivity.png", "healing.png", "contrast.png"):
        assert (tmp_path / fn).stat().st_size > 1000, fn
