"""Field generator properties, the C ABI's exported symbols, the Python
binding's coverage of them, and the driver's command-line contract (no GPU
needed: nothing here launches a kernel)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from mnhack8_amd import build, fields, heat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "heat.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(heat_\w+)\s*\(", src, re.M)))


def test_library_exports_every_header_function():
    out = subprocess.run(["nm", "-D", "--defined-only", build.LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (heat_\w+)", out))
    missing = [f for f in header_functions() if f not in exported]
    assert not missing, missing
    assert len(header_functions()) >= 30


def test_python_binding_declares_every_function():
    assert sorted(heat.SIGNATURES) == header_functions()
    lib = heat.load_library()
    for name in heat.SIGNATURES:
        assert getattr(lib, name) is not None


def test_no_cpu_fallback(tmp_path):
    with pytest.raises(ImportError):
        heat._lib = None
        try:
            heat.load_library(str(tmp_path / "missing.so"))
        finally:
            heat._lib = None


def test_driver_usage_advertises_capabilities():
    r = subprocess.run([build.DRIVER], capture_output=True, text=True)
    assert r.returncode == 1
    for flag in ["--verify", "--kernel", "--scatter", "--fi-rank", "--thr", "--reproject-bc", "--detect"]:
        assert flag in r.stderr


def test_driver_rejects_like_the_reference(tmp_path):
    base = [build.DRIVER, "8", "0.01", "1e-3", str(tmp_path / "f"), str(tmp_path)]
    r = subprocess.run(base + ["--fi", "3", "1", "0", "5"], capture_output=True, text=True)
    assert r.returncode == 1 and "level H needs >= 2 ranks" in r.stderr
    r = subprocess.run(base + ["--fi", "1", "1", "0", "64"], capture_output=True, text=True)
    assert r.returncode == 1 and "BIT must be in 0..63" in r.stderr
    r = subprocess.run(base + ["--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown option --bogus" in r.stderr


def test_counter_normal_is_order_independent_and_normal():
    idx = np.arange(200000, dtype=np.uint64)
    g = fields.counter_normal(5, idx)
    sub = fields.counter_normal(5, idx[::7][::-1])
    assert np.array_equal(sub, g[::7][::-1])
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01
    assert abs(np.mean(g ** 3)) < 0.02 and abs(np.mean(g ** 4) - 3) < 0.05
    assert not np.array_equal(g, fields.counter_normal(6, idx))


def test_iid_lognormal_clip_and_blocks():
    k, r = fields.iid_lognormal(16, sigma=1.0, clip=3.0, seed=1)
    assert k.min() >= np.exp(-3) - 1e-15 and k.max() <= np.exp(3) + 1e-12
    assert (r == 1.0).all()
    kb, _ = fields.block_lognormal(16, sigma=2.0, block=4, seed=2)
    blocks = kb.reshape(4, 4, 4, 4, 4, 4).transpose(0, 2, 4, 1, 3, 5).reshape(64, 64)
    assert (blocks == blocks[:, :1]).all()


def test_upscale_rules():
    fine = np.arange(1, 65, dtype=float).reshape(4, 4, 4)
    a = fields.upscale(fine, 2, "arithmetic")
    h = fields.upscale(fine, 2, "harmonic")
    blk = fine[:2, :2, :2].ravel()
    assert a[0, 0, 0] == blk.mean()
    assert abs(h[0, 0, 0] - 1 / np.mean(1 / blk)) < 1e-14
    assert (h <= a).all()


def test_field_files_round_trip(tmp_path):
    k, r = fields.lognormal(8, sigma=1.0, seed=3)
    meta = fields.write_field(str(tmp_path / "f"), k, r, {"model": "lognormal"})
    k2, r2 = fields.read_field(str(tmp_path / "f"), 8)
    assert np.array_equal(k2, k.ravel()) and np.array_equal(r2, r.ravel())
    m = json.load(open(tmp_path / "f_meta.json"))
    assert m["kmin"] == float(k.min()) and m["n_elements"] == 512 and meta["contrast"] == m["contrast"]


def test_gen_field_cli(tmp_path):
    out = str(tmp_path / "h")
    subprocess.run(["python", os.path.join(ROOT, "mnhack8_amd", "gen_field.py"), "--n", "4", "--model",
                    "homogeneous", "--k", "2.5", "--out", out], check=True, capture_output=True)
    assert np.all(np.fromfile(out + "_k.bin") == 2.5) and np.fromfile(out + "_rhoc.bin").size == 64


def test_seu_campaign_summary_logic():
    """Classification bookkeeping of the upset campaign (host logic only):
    the common population per precision, the refined solve's whole state
    (its FP64 sites added), D1-only coverage and the interval of a
    difference of coverages."""
    from mnhack8_amd import seu_campaign as sc
    rows = []
    def row(prec, pop, vec, cls, d1):
        return {"precision": prec, "population": pop, "vec": vec, "class": cls, "fired": 1, "d1_fired": d1,
                "d2_fired": 0, "nonfinite": 0, "status_fired": 0, "stuck": 0,
                "detected_without_S": int(cls == "detected")}
    for prec in ("fp64", "fp32_refined"):
        for i, cls in enumerate(["benign"] * 6 + ["detected"] * 3 + ["silent"]):
            rows.append(row(prec, "common", sc.VECS[i % 5], cls, int(cls == "detected" and i % 2 == 0)))
    rows += [row("fp32_refined", "fp64_state", "x64", "benign", 0), row("fp32_refined", "fp64_state", "Ax64",
                                                                         "detected", 0)]
    s = sc.summarize(rows, {})
    assert s["fp64"]["benign"] == 6 and s["fp64"]["detected"] == 3 and s["fp64"]["silent"] == 1
    assert s["fp64"]["coverage"] == 0.75
    assert s["fp32_refined"]["sites"] == 10 and s["fp32_refined_whole_state"]["sites"] == 12
    assert s["fp32_refined_whole_state"]["coverage"] == 0.8
    assert abs(s["coverage_difference_pp"] - 5.0) < 1e-12
    lo, hi = s["coverage_difference_ci95_pp"]
    assert lo < 5.0 < hi and -100 <= lo and hi <= 100
    assert s["fp64"]["coverage_D1_only"]["detected"] == 2
    assert s["fp32_refined_whole_state"]["by_vector"]["Ax64"]["detected"] == 1
    lo, hi = s["fp64"]["coverage_ci95"]
    assert 0.2 < lo < 0.75 < hi <= 1.0
    assert sc.wilson(0, 0)[0] != sc.wilson(0, 0)[0]       # NaN for no harmful sites
    # Newcombe's interval of a difference: symmetric under swapping the two
    # proportions (with the sign), and it holds 0 for equal proportions
    l1, h1 = sc.newcombe(7, 20, 12, 25)
    l2, h2 = sc.newcombe(12, 25, 7, 20)
    assert abs(l1 + h2) < 1e-12 and abs(h1 + l2) < 1e-12
    l3, h3 = sc.newcombe(10, 40, 5, 20)
    assert l3 < 0 < h3
