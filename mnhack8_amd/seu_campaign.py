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
  ST  the refined solve reports that its outer residual was not finite (it
      stops there: no correction can recover it, heat_pcg_mixed_status 2).
Classes: benign if the returned iterate meets the solve's own tolerance
(true relative residual <= 10 tol); otherwise detected (any detector) or
silent (none).  Coverage = detected / (detected + silent), with a Wilson 95 %
interval.  Both with and without S are reported, and D1 (the ABFT check)
alone.

Populations.  The drawn sites (vector of {u, r, p, z, Ap}, node, bit
fraction, iteration fraction) are the same for both solves: the same vector
and node, the bit at the same fraction of the word and the iteration at the
same fraction of the clean solve.  The refined solve also keeps FP64 state
of its own, the iterate x and the outer apply's A x (once per outer pass);
a further set of sites, in proportion 2 of 7 arrays, draws those (bits
0..63, outer pass at a fraction of the clean solve's passes), and the
refined solve's coverage is reported over its whole state as well.  The
difference of two coverages carries Newcombe's hybrid-score 95 % interval.
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
VECS64 = ("x64", "Ax64")           # the refined solve's FP64 state


def wilson(k, n, z=1.96):
    if n == 0:
        return (float("nan"), float("nan"))
    p = k / n
    d = 1 + z * z / n
    c = (p + z * z / (2 * n)) / d
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / d
    return (c - h, c + h)


def newcombe(k1, n1, k2, n2):
    """95 % interval of p1 - p2 from the two Wilson intervals (Newcombe's
    hybrid score method, Stat. Med. 17 (1998) 873, method 10)."""
    if n1 == 0 or n2 == 0:
        return (float("nan"), float("nan"))
    p1, p2 = k1 / n1, k2 / n2
    l1, u1 = wilson(k1, n1)
    l2, u2 = wilson(k2, n2)
    d = p1 - p2
    return (d - math.sqrt((p1 - l1) ** 2 + (u2 - p2) ** 2), d + math.sqrt((u1 - p1) ** 2 + (p2 - l2) ** 2))


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
        """(work, stuck, d1, calls, passes, status): iterations (FP64) or
        inner iterations (refined), the limit reached, the D1 record, the
        outer passes and the refined solve's end status (1 for FP64)."""
        op = self.op
        op.copy(self.zero, "T")
        op.abft_take()
        if prec == "fp64":
            it = op.pcg_solve("b", "T", self.max_iter, self.tol)
            stuck = it >= self.max_iter
            work, passes, status = it, 0, 1
        else:
            outer, inner = op.pcg_solve_mixed("b", "T", self.max_outer, self.max_inner, self.tol, self.tol_inner)
            status = op.mixed_status()
            stuck = status == 0
            work, passes = inner, outer
        d1, calls = op.abft_take()
        return work, stuck, d1, calls, passes, status

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
            work, stuck, d1, calls, passes, status = self.solve(prec)
            self.op.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            self.op.copy("T", xref)
            out[prec] = {"iterations": work, "outer_passes": passes, "d1_floor": d1, "d1_calls": calls, "ms": ms,
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
    def site(self, prec, xref, vec, node, bit, it_site, index, population):
        self.op.arm_pcg_fault(vec, node, bit, it_site)
        work, stuck, d1, calls, passes, status = self.solve(prec)
        fired = self.op.pcg_fault_fired()
        fin = self.finite()
        res = self.true_residual() if fin else float("inf")
        err = self.error_vs(xref) if fin else float("inf")
        d1f = (not math.isfinite(d1)) or d1 > self.thr[prec]
        d2f = not (res <= 1e-6)
        stf = status == 2
        benign = fin and res <= 10 * self.tol
        det = (not fin) or d1f or d2f or stf
        cls = "benign" if benign else ("detected" if det or stuck else "silent")
        return {"precision": prec, "population": population, "site": index, "vec": vec, "node": node, "bit": bit,
                "iteration": it_site, "fired": int(fired), "work": work, "passes": passes, "stuck": int(stuck),
                "d1": d1, "d1_fired": int(d1f), "true_residual": res, "d2_fired": int(d2f),
                "nonfinite": int(not fin), "status_fired": int(stf), "E": err, "class": cls,
                "detected_without_S": int(det)}

    def run(self, sites=2000, seed=1, log=None):
        rng = np.random.default_rng(seed)
        draws = [(int(rng.integers(len(VECS))), int(rng.integers(self.nn)), float(rng.random()),
                  float(rng.random())) for _ in range(sites)]
        # the refined solve's own FP64 state: 2 arrays of 7
        n64 = int(round(sites * len(VECS64) / (len(VECS) + len(VECS64))))
        draws64 = [(int(rng.integers(len(VECS64))), int(rng.integers(self.nn)), float(rng.random()),
                    float(rng.random())) for _ in range(n64)]
        rows = []
        for prec, xref, wbits in (("fp64", self.x64, 64), ("fp32_refined", self.x32, 32)):
            iters = self.clean_info[prec]["iterations"]
            t0 = time.perf_counter()
            for s, (v, node, ub, ui) in enumerate(draws):
                bit = min(wbits - 1, int(ub * wbits))
                it_site = min(iters - 1, int(ui * iters))
                rows.append(self.site(prec, xref, VECS[v], node, bit, it_site, s, "common"))
                if log and (s + 1) % 500 == 0:
                    log(f"{prec}: {s + 1}/{sites} sites, {time.perf_counter() - t0:.1f} s")
        passes = max(1, self.clean_info["fp32_refined"]["outer_passes"])
        t0 = time.perf_counter()
        for s, (v, node, ub, ui) in enumerate(draws64):
            bit = min(63, int(ub * 64))
            it_site = min(passes - 1, int(ui * passes))
            rows.append(self.site("fp32_refined", self.x32, VECS64[v], node, bit, it_site, s, "fp64_state"))
        if log:
            log(f"fp32_refined FP64 state: {n64} sites, {time.perf_counter() - t0:.1f} s")
        return rows


def _cov(harm, key=None):
    n = len(harm)
    k = sum(1 for r in harm if (r[key] if key else r["class"] == "detected"))
    return {"harmful": n, "detected": k, "coverage": k / n if n else float("nan"), "ci95": wilson(k, n)}


def _summary(rs):
    harm = [r for r in rs if r["class"] != "benign"]
    det = sum(1 for r in harm if r["class"] == "detected")
    n_h = len(harm)
    by_vec = {}
    for v in VECS + VECS64:
        sv = [r for r in rs if r["vec"] == v]
        if not sv:
            continue
        hv = [r for r in harm if r["vec"] == v]
        by_vec[v] = {"sites": len(sv), **_cov(hv), "silent": sum(1 for r in hv if r["class"] == "silent"),
                     "D1_only": _cov(hv, "d1_fired")}
    return {
        "sites": len(rs), "fired": sum(r["fired"] for r in rs),
        "benign": len(rs) - n_h, "detected": det, "silent": n_h - det,
        "coverage": det / n_h if n_h else float("nan"), "coverage_ci95": wilson(det, n_h),
        "coverage_without_S": (sum(1 for r in harm if r["detected_without_S"]) / n_h) if n_h else float("nan"),
        "coverage_D1_only": _cov(harm, "d1_fired"),
        "by_detector": {"D1": sum(r["d1_fired"] for r in harm), "D2": sum(r["d2_fired"] for r in harm),
                        "NF": sum(r["nonfinite"] for r in harm), "ST": sum(r["status_fired"] for r in harm),
                        "S": sum(r["stuck"] for r in harm)},
        "by_vector": by_vec,
    }


def summarize(rows, clean):
    out = {"clean": clean}
    common = [r for r in rows if r["population"] == "common"]
    out["fp64"] = _summary([r for r in common if r["precision"] == "fp64"])
    out["fp32_refined"] = _summary([r for r in common if r["precision"] == "fp32_refined"])
    out["fp32_refined_whole_state"] = _summary([r for r in rows if r["precision"] == "fp32_refined"])
    a, b = out["fp64"], out["fp32_refined_whole_state"]
    out["coverage_difference_pp"] = 100 * (b["coverage"] - a["coverage"])
    lo, hi = newcombe(b["detected"], b["detected"] + b["silent"], a["detected"], a["detected"] + a["silent"])
    out["coverage_difference_ci95_pp"] = (100 * lo, 100 * hi)
    da, db = a["coverage_D1_only"], b["coverage_D1_only"]
    out["d1_coverage_difference_pp"] = 100 * (db["coverage"] - da["coverage"])
    lo, hi = newcombe(db["detected"], db["harmful"], da["detected"], da["harmful"])
    out["d1_coverage_difference_ci95_pp"] = (100 * lo, 100 * hi)
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
    print(json.dumps({k: summ[k] for k in ("fp64", "fp32_refined", "fp32_refined_whole_state",
                                           "coverage_difference_pp", "coverage_difference_ci95_pp",
                                           "d1_coverage_difference_pp", "d1_coverage_difference_ci95_pp")}),
          flush=True)


if __name__ == "__main__":
    main()
