"""Single-event-upset campaign over the PCG state of one backward-Euler solve.

Sites are drawn from a seed uniformly over vector (u, r, p, z, Ap) x node x
bit x iteration, for the FP64 PCG (64-bit words, iterations of the clean FP64
solve) and for the FP32-refined solve (the FP32 inner state: 32-bit words,
inner iterations of the clean refined solve).  Each site is one solve with a
one-shot flip of the stored copy of that vector where that iteration
produces it (HeatOperator.arm_pcg_fault).

Detectors, as in the reference's run_case (src/heat_solver_het.cu:1202-1234):
  D1  the operator checksum of every apply in the solve: max rel > threshold.
      FP64: chk_tol = 1e-10 (src/heat_solver_het.cu:637).  FP32 inner applies:
      the threshold is re-derived from the FP32 noise floor, 10 x the largest
      clean rel over calibration solves (the reference's own 10 x floor rule
      for E, src/heat_solver_het.cu:1261).
  D2  the true residual of the returned iterate: ||b - A x|| / ||b|| > 1e-6
      (src/heat_solver_het.cu:1202-1210).
  NF  a non-finite iterate.
  S   the solver stops without reaching its tolerance (iteration limit).
Classes: benign if the returned iterate meets the solve's own tolerance
(true relative residual <= 10 tol); otherwise detected (any detector) or
silent (none).  Coverage = detected / (detected + silent), with a Wilson 95 %
interval.  Both with and without S are reported.
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import os
import sys
import time

import numpy as np

VECS = ("u", "r", "p", "z", "Ap")


def wilson(k, n, z=1.96):
    if n == 0:
        return (float("nan"), float("nan"))
    p = k / n
    d = 1 + z * z / n
    c = (p + z * z / (2 * n)) / d
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / d
    return (c - h, c + h)


class Campaign:
    def __init__(self, n=128, dt=1e-4, seed=1, tol=1e-10, tol_inner=1e-4, max_iter=500, max_outer=50,
                 max_inner=500):
        import torch
        from . import fields
        from .heat import HeatOperator
        self.torch = torch
        self.n, self.dt, self.tol, self.tol_inner = n, dt, tol, tol_inner
        self.max_iter, self.max_outer, self.max_inner = max_iter, max_outer, max_inner
        k, rhoc = fields.iid_lognormal(n, sigma=1.0, clip=3.0, seed=seed)
        self.op = HeatOperator(n, dt, k.ravel(), rhoc.ravel(), kernel="fast", scatter="colored", detect=True,
                               source="const", source_value=1.0)
        op = self.op
        # an armed upset of the refined solve runs its standard inner
        # iteration; the clean runs and floors use the same one
        op.set_inner_solver("standard")
        self.nn = op.nn
        dev = lambda: torch.zeros(self.nn, dtype=torch.float64, device="cuda")
        self.zero, self.tmp, self.x64, self.x32, self.bsave = dev(), dev(), dev(), dev(), dev()
        op.upload("T", np.zeros(self.nn))
        op.build_rhs("T", "b")
        op.copy("b", self.bsave)
        op.synchronize()
        self.bnorm = op.norm2("b")

    # -- one solve -------------------------------------------------------
    def solve(self, prec):
        op = self.op
        op.copy(self.zero, "T")
        op.abft_take()
        if prec == "fp64":
            it = op.pcg_solve("b", "T", self.max_iter, self.tol)
            stuck = it >= self.max_iter
            work = it
        else:
            outer, inner = op.pcg_solve_mixed("b", "T", self.max_outer, self.max_inner, self.tol, self.tol_inner)
            stuck = outer >= self.max_outer
            work = inner
        d1, calls = op.abft_take()
        return work, stuck, d1, calls

    def true_residual(self):
        op = self.op
        op.apply_operator("T", self.tmp, self.dt)
        op.axpy(-1.0, self.bsave, self.tmp)          # tmp = A x - b
        return op.norm2(self.tmp) / self.bnorm

    def error_vs(self, xref):
        op = self.op
        op.copy("T", self.tmp)
        op.axpy(-1.0, xref, self.tmp)
        den = op.norm2(xref)
        return op.norm2(self.tmp) / den if den > 0 else float("nan")

    def finite(self):
        # min/max skip NaN (fmin/fmax, as the reference's D3), the norm does not
        lo, hi = self.op.minmax("T")
        return math.isfinite(lo) and math.isfinite(hi) and not math.isnan(self.op.norm2("T"))

    # -- calibration and clean runs --------------------------------------
    def clean(self):
        out = {}
        for prec, xref in (("fp64", self.x64), ("fp32_refined", self.x32)):
            t0 = time.perf_counter()
            work, stuck, d1, calls = self.solve(prec)
            self.op.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            self.op.copy("T", xref)
            out[prec] = {"iterations": work, "d1_floor": d1, "d1_calls": calls, "ms": ms,
                         "true_residual": self.true_residual()}
        # FP32 D1 noise floor over further right-hand sides (later time steps)
        floors = [out["fp32_refined"]["d1_floor"]]
        op = self.op
        op.copy(self.x32, "T")
        for step in range(3):
            op.build_rhs("T", "b")
            op.copy("b", self.tmp)
            op.abft_take()
            op.pcg_solve_mixed("b", "T", self.max_outer, self.max_inner, self.tol, self.tol_inner)
            floors.append(op.abft_take()[0])
        op.copy(self.bsave, "b")
        self.thr = {"fp64": op.abft_threshold(), "fp32_refined": 10.0 * max(floors)}
        out["fp32_refined"]["d1_floor_steps"] = floors
        out["thresholds"] = self.thr
        self.clean_info = out
        return out

    # -- the campaign ----------------------------------------------------
    def run(self, sites=2000, seed=1, log=None):
        rng = np.random.default_rng(seed)
        draws = [(int(rng.integers(len(VECS))), int(rng.integers(self.nn)), float(rng.random()),
                  float(rng.random())) for _ in range(sites)]
        rows = []
        for prec, xref, wbits in (("fp64", self.x64, 64), ("fp32_refined", self.x32, 32)):
            iters = self.clean_info[prec]["iterations"]
            t0 = time.perf_counter()
            for s, (v, node, ub, ui) in enumerate(draws):
                bit = min(wbits - 1, int(ub * wbits))
                it_site = min(iters - 1, int(ui * iters))
                self.op.arm_pcg_fault(VECS[v], node, bit, it_site)
                work, stuck, d1, calls = self.solve(prec)
                fired = self.op.pcg_fault_fired()
                fin = self.finite()
                res = self.true_residual() if fin else float("inf")
                err = self.error_vs(xref) if fin else float("inf")
                d1f = (not math.isfinite(d1)) or d1 > self.thr[prec]
                d2f = not (res <= 1e-6)
                benign = fin and res <= 10 * self.tol
                det = (not fin) or d1f or d2f
                cls = "benign" if benign else ("detected" if det or stuck else "silent")
                rows.append({"precision": prec, "site": s, "vec": VECS[v], "node": node, "bit": bit,
                             "iteration": it_site, "fired": int(fired), "work": work, "stuck": int(stuck),
                             "d1": d1, "d1_fired": int(d1f), "true_residual": res, "d2_fired": int(d2f),
                             "nonfinite": int(not fin), "E": err, "class": cls,
                             "detected_without_S": int(det)})
                if log and (s + 1) % 250 == 0:
                    log(f"{prec}: {s + 1}/{sites} sites, {time.perf_counter() - t0:.1f} s")
        return rows


def summarize(rows, clean):
    out = {"clean": clean}
    for prec in ("fp64", "fp32_refined"):
        rs = [r for r in rows if r["precision"] == prec]
        harm = [r for r in rs if r["class"] != "benign"]
        det = sum(1 for r in harm if r["class"] == "detected")
        det_noS = sum(1 for r in harm if r["detected_without_S"])
        n_h = len(harm)
        by_vec = {}
        for v in VECS:
            hv = [r for r in harm if r["vec"] == v]
            by_vec[v] = {"sites": sum(1 for r in rs if r["vec"] == v), "harmful": len(hv),
                         "detected": sum(1 for r in hv if r["class"] == "detected"),
                         "silent": sum(1 for r in hv if r["class"] == "silent")}
        out[prec] = {
            "sites": len(rs), "fired": sum(r["fired"] for r in rs),
            "benign": len(rs) - n_h, "detected": det, "silent": n_h - det,
            "coverage": det / n_h if n_h else float("nan"), "coverage_ci95": wilson(det, n_h),
            "coverage_without_S": det_noS / n_h if n_h else float("nan"),
            "coverage_D1_only": sum(r["d1_fired"] for r in harm) / n_h if n_h else float("nan"),
            "by_detector": {"D1": sum(r["d1_fired"] for r in harm), "D2": sum(r["d2_fired"] for r in harm),
                            "NF": sum(r["nonfinite"] for r in harm), "S": sum(r["stuck"] for r in harm)},
            "by_vector": by_vec,
        }
    a, b = out["fp64"]["coverage"], out["fp32_refined"]["coverage"]
    out["coverage_difference_pp"] = 100 * (b - a)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--sites", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default="seu_campaign")
    a = ap.parse_args(argv)
    log = lambda m: print(f"[seu] {m}", file=sys.stderr, flush=True)
    os.makedirs(a.out, exist_ok=True)
    c = Campaign(n=a.n, seed=a.seed)
    clean = c.clean()
    log(f"clean: {json.dumps(clean)}")
    rows = c.run(a.sites, seed=a.seed, log=log)
    with open(os.path.join(a.out, "sites.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    summ = summarize(rows, clean)
    with open(os.path.join(a.out, "summary.json"), "w") as f:
        json.dump(summ, f, indent=1)
    print(json.dumps({k: summ[k] for k in ("fp64", "fp32_refined", "coverage_difference_pp")}), flush=True)


if __name__ == "__main__":
    main()
