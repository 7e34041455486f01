#!/usr/bin/env python3
"""Acceptance gates V0-V6 for a solver binary with the heat_solver_het CLI.

Restates the gates of the reference's validation stages
(scripts/validations/v0_convergence.py, v1_regression.py, v2_noise_floor.py,
v3_injection.py, v4_v5_equivalence.py, v6_detection.py) so that they can run
against this solver's driver (mnhack8_amd/_build/heat_solver_het) or any
binary with the same command line.  Each stage writes <out>/<stage>/report.json
({test, pass, details}, as scripts/validations/vlib.py:43-49) and the run
writes <out>/summary.txt.

    python -m mnhack8_amd.validations --bin BIN --out DIR [--stages v0,v1,v2,v3,v5,v6]

V4 (np = 2, 4 against np = 1) needs several ranks; this solver owns the whole
mesh on one GPU, so V4 reports UNSUPPORTED.  V1 compares with the MNHack7
prototype's printed norms recorded by the reference (the prototype is not
rebuilt here).
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import os
import subprocess
import sys

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from mnhack8_amd import fields  # noqa: E402

# MNHack7 prototype, quadrants field, t = 0.01, dt = 1e-4: final ||T||_2 as
# printed with 7 digits (tests/baseline_validation/v10/v1/report.json:7, 58).
PROTOTYPE_NORM = {16: 0.2836269, 15: 0.2579253}


def run(cmd, log=None, env=None):
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True, env=env)
    if log:
        with open(log, "a") as f:
            f.write(f"$ {' '.join(map(str, cmd))}\n[rc={r.returncode}]\n{r.stdout}\n{r.stderr}\n")
    if r.returncode != 0:
        raise RuntimeError(f"command failed (rc={r.returncode}): {' '.join(map(str, cmd))}\n{r.stderr[-2000:]}")
    return r.stdout


def capabilities(binary):
    """Flags advertised in the usage text (scripts/validations/vlib.py:29-33)."""
    r = subprocess.run([binary], capture_output=True, text=True)
    txt = r.stdout + r.stderr
    return {f for f in ["--verify", "--kernel", "--scatter", "--fi-rank", "--thr", "--reproject-bc", "--detect"]
            if f in txt}


def read_csv(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def fnum(x):
    try:
        return float(x)
    except (TypeError, ValueError):
        return float("nan")


def write_report(outdir, name, passed, details):
    os.makedirs(outdir, exist_ok=True)
    rep = {"test": name, "pass": bool(passed), "details": details}
    with open(os.path.join(outdir, "report.json"), "w") as f:
        json.dump(rep, f, indent=2, default=float)
    print(f"[{name}] {'PASS' if passed else 'FAIL'}", flush=True)
    return rep


def gen(model, n, out, **kw):
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    if not os.path.exists(out + "_k.bin"):
        if model == "homogeneous":
            k, r = fields.homogeneous(n, kw.get("k", 1.0), kw.get("rhoc", 1.0))
        elif model == "quadrants":
            k, r = fields.quadrants(n)
        else:
            k, r = fields.lognormal(n, sigma=kw.get("sigma", 1.0), L=kw.get("L", 0.125), seed=kw.get("seed", 7))
        fields.write_field(out, k, r, {"n": n, "model": model, **kw})
    return out


def node_index(ix, iy, iz, N):
    return (ix * (N + 1) + iy) * (N + 1) + iz


def elem_index(ix, iy, iz, N):
    return (ix * N + iy) * N + iz


# --------------------------------------------------------------------------
# V0 eigenmode convergence (scripts/validations/v0_convergence.py:14-31, 73-90)
SPATIAL = [(8, 0.02, 1e-5), (16, 0.02, 1e-5), (32, 0.02, 1e-5)]
TEMPORAL = [(32, 0.04, 4e-3), (32, 0.04, 2e-3), (32, 0.04, 1e-3), (32, 0.04, 5e-4)]
LAM = 3 * math.pi ** 2


def v0(binary, out, kernel="fast", scatter="atomic", extra=()):
    d = os.path.join(out, f"v0_{kernel}_{scatter}_np1")
    os.makedirs(d, exist_ok=True)
    vpath = os.path.join(d, "verify.csv")
    if os.path.exists(vpath):
        os.remove(vpath)
    for N, t, dt in SPATIAL + TEMPORAL:
        fp = gen("homogeneous", N, os.path.join(d, f"homog{N}"), k=1.0, rhoc=1.0)
        run([binary, N, t, dt, fp, d, "--verify", "--kernel", kernel, "--scatter", scatter, *extra],
            log=os.path.join(d, "run.log"))
    err = {(int(r["N"]), float(r["dt"])): float(r["rel_L2_err"]) for r in read_csv(vpath)}
    det = {"spatial": [], "temporal": []}
    ok = True
    es = [err[(N, dt)] for N, _, dt in SPATIAL]
    for (N, t, dt), e in zip(SPATIAL, es):
        pred = math.pi ** 4 * (1.0 / N) ** 2 * t / 4
        det["spatial"].append({"N": N, "err": e, "pred": pred, "ratio_to_pred": e / pred})
        ok &= 0.5 <= e / pred <= 2.0
    orders = [math.log2(es[i] / es[i + 1]) for i in range(len(es) - 1)]
    det["spatial_orders"] = orders
    ok &= 1.7 <= orders[0] <= 2.3 and orders[1] >= 1.6
    et = [err[(N, dt)] for N, _, dt in TEMPORAL]
    for (N, t, dt), e in zip(TEMPORAL, et):
        pred = LAM ** 2 * dt * t / 2
        det["temporal"].append({"dt": dt, "err": e, "pred": pred, "ratio_to_pred": e / pred})
        ok &= 0.5 <= e / pred <= 2.0
    torders = [math.log2(et[i] / et[i + 1]) for i in range(len(et) - 1)]
    det["temporal_orders"] = torders
    ok &= all(0.85 <= o <= 1.15 for o in torders)
    return write_report(d, "v0", ok, det)


# --------------------------------------------------------------------------
# V1 regression to the MNHack7 prototype (scripts/validations/v1_regression.py:36-64)
def v1(binary, out, t=0.01, dt=1e-4, extra=()):
    d0 = os.path.join(out, "v1")
    res = {}
    for N in (16, 15):
        d = os.path.join(d0, f"v2_N{N}")
        os.makedirs(d, exist_ok=True)
        fp = gen("quadrants", N, os.path.join(d0, f"quad{N}"))
        run([binary, N, t, dt, fp, d, "--tag", "reg", "--snap", "0", "--kernel", "gauss", "--scatter", "atomic",
             *extra], log=os.path.join(d0, "run.log"))
        T = np.fromfile(os.path.join(d, "T_final_reg_clean_rank0.bin"))
        v2_norm = float(np.linalg.norm(T))
        res[N] = {"N": N, "ref_norm": PROTOTYPE_NORM[N], "v2_norm": v2_norm,
                  "rel": abs(v2_norm - PROTOTYPE_NORM[N]) / PROTOTYPE_NORM[N]}
    ok = res[16]["rel"] <= 1e-6 and res[15]["rel"] > 1e-6
    return write_report(d0, "v1", ok, {"even": res[16], "odd_negative_control": res[15]})


# --------------------------------------------------------------------------
# V2 noise floor (scripts/validations/v2_noise_floor.py:25-29, 48-67)
def v2(binary, out, K=20, N=32, t=0.02, dt=5e-4, extra=()):
    d0 = os.path.join(out, "v2")
    fp = gen("lognormal", N, os.path.join(d0, f"ln{N}"), sigma=1.0, L=0.125, seed=7)
    det = {}
    ok = True
    for scatter in ("colored", "atomic"):
        vals, finals = [], []
        for k in range(K):
            d = os.path.join(d0, f"{scatter}_{k:02d}")
            os.makedirs(d, exist_ok=True)
            run([binary, N, t, dt, fp, d, "--tag", "nf", "--snap", "1", "--kernel", "fast", "--scatter", scatter,
                 *extra], log=os.path.join(d0, "run.log"))
            sm = read_csv(os.path.join(d, "fi_summary_nf.csv"))[0]
            vals.append(fnum(sm["E_floor_max"]))
            finals.append(np.fromfile(os.path.join(d, "T_final_nf_clean_rank0.bin")))
        vals = np.array(vals)
        bitwise = all(np.array_equal(f.view(np.uint64), finals[0].view(np.uint64)) for f in finals)
        rec = {"E_floor_max_samples": vals.tolist(), "max": float(vals.max()), "mean": float(vals.mean()),
               "all_zero": bool((vals == 0).all()), "cleanA_bitwise_identical_across_invocations": bitwise,
               "thr_recommended": float(max(1e-13, 10 * vals.max()))}
        det[scatter] = rec
        if scatter == "colored":
            ok &= bool((vals == 0).all()) and bitwise
    return write_report(d0, "v2", ok, det)


# --------------------------------------------------------------------------
# V3 injector (scripts/validations/v3_injection.py:51-139), single rank
def _fi_run(binary, N, t, dt, fp, d, level, S, target, bit, tag, extra=(), detect=False):
    os.makedirs(d, exist_ok=True)
    cmd = [binary, N, t, dt, fp, d, "--tag", tag, "--snap", "1", "--kernel", "fast", "--scatter", "colored",
           "--fi", level, S, target, bit, *extra]
    if detect:
        cmd.append("--detect")
    run(cmd, log=os.path.join(os.path.dirname(d), "run.log"))
    tr = read_csv(os.path.join(d, f"fi_trace_{tag}.csv"))
    sm = read_csv(os.path.join(d, f"fi_summary_{tag}.csv"))[0]
    E = np.array([fnum(r["E_fault"]) for r in tr])
    st = np.array([int(r["step"]) for r in tr])
    return E, st, sm


def v3(binary, out, N=32, t=0.02, dt=5e-4, S=10, bits=(0, 30, 51, 55, 62, 63), extra=()):
    d0 = os.path.join(out, "v3")
    fp = gen("lognormal", N, os.path.join(d0, f"ln{N}"), sigma=1.0, L=0.125, seed=7)
    tgt = {1: node_index(N // 2, N // 2, N // 2, N), 2: elem_index(N // 2, N // 2, N // 2, N)}
    det = {"runs": [], "H": "skipped (needs >= 2 ranks)"}
    ok = True
    for level in (1, 2):
        for bit in bits:
            E, st, sm = _fi_run(binary, N, t, dt, fp, os.path.join(d0, f"L{level}_b{bit}"), level, S, tgt[level],
                                bit, "fi", extra)
            before = E[st < S]
            atS = E[st == S]
            c1 = bool((before == 0).all())
            c2 = bool(atS.size == 1 and (atS[0] > 0 or not np.isfinite(atS[0])))
            det["runs"].append(dict(level=level, bit=bit, before_zero=c1, fires_at_S=c2,
                                    E_max=fnum(sm["E_fault_max"]), E_final=fnum(sm["E_fault_final"]),
                                    cls=sm["class"], on_boundary=int(sm["target_on_boundary"]),
                                    iters_delta=int(sm["iters_delta_max"])))
            ok &= c1 and c2
    bnode = node_index(0, N // 2, N // 2, N)
    _, _, sm = _fi_run(binary, N, t, dt, fp, os.path.join(d0, "G_boundary"), 1, S, bnode, 62, "fi", extra)
    _, _, smr = _fi_run(binary, N, t, dt, fp, os.path.join(d0, "G_boundary_reproject"), 1, S, bnode, 62, "fi",
                        ("--reproject-bc", *extra))
    red = fnum(sm["E_fault_final"]) / max(fnum(smr["E_fault_final"]), 1e-300)
    det["boundary_control"] = {"on_boundary_flag": int(sm["target_on_boundary"]),
                               "no_reproject": {"E_max": fnum(sm["E_fault_max"]), "E_final": fnum(sm["E_fault_final"]),
                                                "class": sm["class"]},
                               "reproject": {"E_max": fnum(smr["E_fault_max"]), "E_final": fnum(smr["E_fault_final"]),
                                             "class": smr["class"]},
                               "E_final_reduction_factor": red}
    ok &= int(sm["target_on_boundary"]) == 1 and sm["class"] == "persistent" and red >= 100
    return write_report(d0, "v3", ok, det)


# --------------------------------------------------------------------------
# V5 gauss vs fast (scripts/validations/v4_v5_equivalence.py:76-83); V4 needs ranks
def v5(binary, out, N=16, t=0.02, dt=5e-4, extra=()):
    d0 = os.path.join(out, "v5")
    fp = gen("lognormal", N, os.path.join(out, f"ln{N}"), sigma=1.0, L=0.125, seed=7)
    T, sm = {}, {}
    for kern in ("gauss", "fast"):
        d = os.path.join(d0, kern)
        os.makedirs(d, exist_ok=True)
        run([binary, N, t, dt, fp, d, "--tag", "eq", "--snap", "0", "--kernel", kern, "--scatter", "colored", *extra],
            log=os.path.join(d0, "run.log"))
        T[kern] = np.fromfile(os.path.join(d, "T_final_eq_clean_rank0.bin"))
        sm[kern] = read_csv(os.path.join(d, "fi_summary_eq.csv"))[0]
    r = float(np.linalg.norm(T["fast"] - T["gauss"]) / np.linalg.norm(T["gauss"]))
    det = {"field_rel_diff": r, "solve_ms_gauss": fnum(sm["gauss"]["solve_ms_A"]),
           "solve_ms_fast": fnum(sm["fast"]["solve_ms_A"]),
           "speedup": fnum(sm["gauss"]["solve_ms_A"]) / max(fnum(sm["fast"]["solve_ms_A"]), 1e-9),
           "mean_cg_iters": {"gauss": fnum(sm["gauss"]["mean_cg_iters_A"]),
                             "fast": fnum(sm["fast"]["mean_cg_iters_A"])}}
    return write_report(d0, "v5", r <= 1e-12, det)


def v4(out):
    return write_report(os.path.join(out, "v4"), "v4", False,
                        {"status": "UNSUPPORTED: one rank owns the whole mesh (no np=2,4 runs)"})


# --------------------------------------------------------------------------
# V6 detectors (scripts/validations/v6_detection.py:63-104), single rank
def v6(binary, out, N=32, t=0.02, dt=5e-4, S=10, bits=(30, 51, 55, 62, 63), extra=()):
    d0 = os.path.join(out, "v6")
    fp = gen("lognormal", N, os.path.join(d0, f"ln{N}"), sigma=1.0, L=0.125, seed=7)
    det = {"clean": {}, "coverage": [], "H": "skipped"}
    ok = True
    for scatter in ("colored", "atomic"):
        d = os.path.join(d0, f"clean_{scatter}")
        os.makedirs(d, exist_ok=True)
        run([binary, N, t, dt, fp, d, "--tag", "det", "--snap", "1", "--kernel", "fast", "--scatter", scatter,
             "--detect", *extra], log=os.path.join(d0, "run.log"))
        sm = read_csv(os.path.join(d, "fi_summary_det.csv"))[0]
        rec = {k: fnum(sm[k]) for k in ("d1_max_A", "d2_max_A")}
        rec.update({k: int(sm[k]) for k in ("d1_first", "d2_first", "d3_first")})
        det["clean"][scatter] = rec
        ok &= rec["d1_first"] < 0 and rec["d2_first"] < 0 and rec["d3_first"] < 0
    tgt = {1: node_index(N // 2, N // 2, N // 2, N), 2: elem_index(N // 2, N // 2, N // 2, N)}
    rows = []
    for level in (1, 2):
        for bit in bits:
            _, _, sm = _fi_run(binary, N, t, dt, fp, os.path.join(d0, f"L{level}_b{bit}"), level, S, tgt[level], bit,
                               "det", extra, detect=True)
            r = dict(level="GR"[level - 1], bit=bit, cls=sm["class"], by=sm["detected_by"],
                     d1_first=int(sm["d1_first"]), d2_first=int(sm["d2_first"]), d3_first=int(sm["d3_first"]),
                     E_max=fnum(sm["E_fault_max"]), iters_delta=int(sm["iters_delta_max"]))
            rows.append(r)
            if level == 2 and bit >= 51:
                ok &= r["d1_first"] == S
            if level == 2 and bit == 63:
                ok &= r["d3_first"] == S
            if level == 2 and bit == 62:
                ok &= r["cls"] == "detected" or r["by"] != "none"
    det["coverage"] = rows
    with open(os.path.join(d0, "coverage.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    return write_report(d0, "v6", ok, det)


STAGES = {"v0": lambda b, o, x: [v0(b, o, k, s, x) for k in ("fast", "gauss") for s in ("colored", "atomic")],
          "v1": lambda b, o, x: v1(b, o, extra=x), "v2": lambda b, o, x: v2(b, o, extra=x),
          "v3": lambda b, o, x: v3(b, o, extra=x), "v4": lambda b, o, x: v4(o),
          "v5": lambda b, o, x: v5(b, o, extra=x), "v6": lambda b, o, x: v6(b, o, extra=x)}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--bin", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--stages", default="v0,v1,v2,v3,v4,v5,v6")
    ap.add_argument("--extra", default="", help="extra flags passed to every solver run")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    extra = tuple(a.extra.split()) if a.extra else ()
    for st in a.stages.split(","):
        STAGES[st](a.bin, a.out, extra)
    lines = []
    for root, _, files in sorted(os.walk(a.out)):
        if "report.json" in files:
            rep = json.load(open(os.path.join(root, "report.json")))
            lines.append(f"{os.path.join(root, 'report.json'):70s} {'PASS' if rep['pass'] else 'FAIL'}")
    with open(os.path.join(a.out, "summary.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
