"""Sixteen right-hand sides on one mesh: the multi-RHS workload.

Mesh N=320, kappa log-normal (sigma = 2) constant on 4^3-element blocks,
rho*c = 1, dt = 1e-4; sixteen Gaussian-bump sources with centres uniform in
[0.2, 0.8]^3 and widths uniform in [0.05, 0.2].  The hot kernel is the
multi-RHS operator apply, whose element contraction runs on the matrix
cores (HeatOperator.apply_mrhs).
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np

NRHS = 16


def workload_fields(n=320, seed=1):
    from . import fields
    return fields.block_lognormal(n, sigma=2.0, block=4, seed=seed)


def bumps(n, count=NRHS, seed=1):
    """(nn, count) FP32 node values of Gaussian bumps exp(-|x - c|^2 / (2 w^2))."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.2, 0.8, size=(count, 3))
    w = rng.uniform(0.05, 0.2, size=count)
    t = np.arange(n + 1) / n
    out = np.empty(((n + 1) ** 3, count), dtype=np.float32)
    for r in range(count):
        gx = np.exp(-(t - c[r, 0]) ** 2 / (2 * w[r] ** 2))
        gy = np.exp(-(t - c[r, 1]) ** 2 / (2 * w[r] ** 2))
        gz = np.exp(-(t - c[r, 2]) ** 2 / (2 * w[r] ** 2))
        out[:, r] = np.einsum("i,j,k->ijk", gx, gy, gz).astype(np.float32).ravel()
    return out


def bench_apply(n=320, reps=20, log=None):
    """Time the 16-RHS apply against sixteen single-vector FP32 and FP64 applies."""
    import torch
    from .heat import HeatOperator
    k, rhoc = workload_fields(n)
    dt = 1e-4
    op = HeatOperator(n, dt, k.ravel(), rhoc.ravel(), kernel="fast", scatter="colored", source="const",
                      source_value=1.0)
    nn, ne = op.nn, op.ne
    X = torch.from_numpy(bumps(n)).cuda()
    Y = torch.zeros_like(X)
    torch.cuda.synchronize()
    ms = op.time_mrhs(X, Y, reps)
    # one FP32 and one FP64 single-vector apply, each timed on its own
    x32 = X[:, 0].contiguous()
    y32 = torch.zeros_like(x32)
    op.apply_operator_f32(x32, y32, dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        op.apply_operator_f32(x32, y32, dt)
    op.synchronize()
    ms32 = (time.perf_counter() - t0) * 1e3 / reps
    ms64, _ = op.time_matvec("p", "Ap", reps)
    alg_bytes = 4 * (2 * NRHS * nn + 2 * ne)
    flops = 2 * 16 * 8 * NRHS * ne                   # the 16 x 8 contraction per element and RHS column
    # issued on the matrix cores: one 16 x 16 x 32 bfloat16 product per element
    # (K = 8 nodes x (hi, mid, lo, 0) terms of the exact split)
    mfma_flops = 2 * 16 * 16 * 32 * ne
    out = {
        "n": n, "nodes": nn, "elements": ne, "rhs": NRHS,
        "ms_per_apply": round(ms, 4),
        "gdof_rhs_per_s": round(nn * NRHS / (ms * 1e-3) / 1e9, 2),
        "algorithmic_bytes": alg_bytes,
        "achieved_GBs": round(alg_bytes / (ms * 1e-3) / 1e9, 1),
        "hbm_frac": round(alg_bytes / (ms * 1e-3) / 8e12, 4),
        "contraction_tflops": round(flops / (ms * 1e-3) / 1e12, 2),
        "mfma_bf16_tflops": round(mfma_flops / (ms * 1e-3) / 1e12, 2),
        "mfma_frac_bf16": round(mfma_flops / (ms * 1e-3) / 2.5e15, 4),
        "single_fp32_ms_x16": round(16 * ms32, 3),
        "single_fp64_ms_x16": round(16 * ms64, 3),
        "speedup_vs_16_fp32_applies": round(16 * ms32 / ms, 2),
    }
    if log:
        log(f"16-RHS apply at N={n}: {ms:.3f} ms ({out['gdof_rhs_per_s']} GDOF-RHS/s, "
            f"{out['contraction_tflops']} TF/s of contraction, {out['mfma_bf16_tflops']} TF/s bf16 issued), "
            f"16 FP32 applies {16 * ms32:.3f} ms, 16 FP64 applies {16 * ms64:.3f} ms")
    op.cleanup()
    return out


def rounding(n=96, seed=7):
    """Rounding of the 16-RHS apply against the FP64 single-vector apply of
    each column on the workload fields at mesh n: eight bump columns and
    eight standard-normal columns; per node, the error relative to the
    column's largest |y|.  Returns max, rms and mean per column kind."""
    import torch
    from .heat import HeatOperator
    k, rhoc = workload_fields(n)
    dt = 1e-4
    op = HeatOperator(n, dt, k.ravel(), rhoc.ravel(), kernel="fast", scatter="colored", source="const",
                      source_value=1.0)
    rng = np.random.default_rng(seed)
    X = np.concatenate([bumps(n)[:, :8], rng.standard_normal((op.nn, 8)).astype(np.float32)], axis=1)
    Xd = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    Yd = torch.zeros_like(Xd)
    op.apply_mrhs(Xd, Yd, dt)
    torch.cuda.synchronize()
    Y = Yd.cpu().numpy().astype(np.float64)
    out = {"n": n}
    for name, cols in (("bumps", range(8)), ("normal", range(8, 16))):
        mx, rms, mean = [], [], []
        for r in cols:
            x = torch.from_numpy(X[:, r].astype(np.float64)).cuda()
            y = torch.zeros_like(x)
            op.apply_operator(x, y, dt)
            torch.cuda.synchronize()
            y64 = y.cpu().numpy()
            d = (Y[:, r] - y64) / np.abs(y64).max()
            mx.append(float(np.abs(d).max()))
            rms.append(float(np.sqrt((d ** 2).mean())))
            mean.append(float(d.mean()))
        out[name] = {"max": max(mx), "rms": float(np.mean(rms)), "mean": float(np.mean(mean))}
    op.cleanup()
    return out


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=320)
    ap.add_argument("--solve-steps", type=int, default=0, help="also time this many steps of the 16-RHS solve")
    ap.add_argument("--refined-steps", type=int, default=0,
                    help="also time this many steps of the 16-column refined solve (matrix-core operator)")
    ap.add_argument("--max-outer", type=int, default=12)
    ap.add_argument("--rounding", type=int, default=0, metavar="N",
                    help="also report the apply's rounding against the FP64 apply at mesh N")
    a = ap.parse_args(argv)
    log = lambda m: print(f"[mrhs] {m}", file=sys.stderr, flush=True)
    out = {"apply": bench_apply(a.n, log=log)}
    if a.rounding:
        out["rounding"] = rounding(a.rounding)
        log(f"rounding at N={a.rounding}: {out['rounding']}")
    if a.solve_steps:
        out["solve"] = bench_solve(a.n, a.solve_steps, log=log)
    if a.refined_steps:
        out["solve_refined"] = bench_solve_refined(a.n, a.refined_steps, max_outer=a.max_outer, log=log)
    print(json.dumps(out))



def load_vectors(n, count=NRHS, seed=1):
    """(count, nn) nodal loads h^3 f_r(x_i) of the Gaussian bumps, zero on the boundary."""
    X = bumps(n, count, seed).astype(np.float64).T.copy()
    h3 = (1.0 / n) ** 3
    t = np.arange(n + 1)
    ix, iy, iz = np.meshgrid(t, t, t, indexing="ij")
    interior = ((ix > 0) & (ix < n) & (iy > 0) & (iy < n) & (iz > 0) & (iz < n)).ravel()
    X *= h3
    X[:, ~interior] = 0.0
    return X


def bench_solve(n=320, steps=20, tol=1e-8, max_iter=500, log=None):
    """Sixteen right-hand sides on one mesh advanced together (shared-mesh
    batch, fused FP64 PCG), against the same sixteen one after another."""
    from .heat import HeatBatch, HeatOperator
    k, rhoc = workload_fields(n)
    dt = 1e-4
    op = HeatOperator(n, dt, k.ravel(), rhoc.ravel(), kernel="fast", scatter="colored", source="const",
                      source_value=1.0)
    F = load_vectors(n)
    t0 = time.perf_counter()
    batch = HeatBatch.shared(op, F)
    setup_s = time.perf_counter() - t0
    zero = np.zeros(op.nn)
    its = []
    batch_s = 0.0
    for st in range(steps):                                # one step per call: progress lines
        t0 = time.perf_counter()
        its.append(batch.step(1, max_iter, tol)[0])
        batch_s += time.perf_counter() - t0
        if log:
            log(f"batched step {st + 1}/{steps}: {(time.perf_counter() - t0) * 1e3:.0f} ms, "
                f"iterations max {int(its[-1].max())} min {int(its[-1].min())}")
    its = np.array(its)
    finals = [op.download(batch.vector(i, "T")) for i in range(NRHS)]
    batch.cleanup()
    # the same right-hand sides one after another (one-member batches on the same mesh)
    seq_s, same = 0.0, True
    for i in range(NRHS):
        one = HeatBatch.shared(op, F[i:i + 1])
        op.upload(one.vector(0, "T"), zero)
        op.synchronize()
        t0 = time.perf_counter()
        for st in range(steps):
            one.step(1, max_iter, tol)
        seq_s += time.perf_counter() - t0
        same = same and np.array_equal(op.download(one.vector(0, "T")), finals[i])
        one.cleanup()
        if log:
            log(f"one after another: right-hand side {i + 1}/{NRHS} done ({seq_s * 1e3:.0f} ms so far)")
    out = {"n": n, "rhs": NRHS, "steps": steps, "tol": tol, "nodes": op.nn,
           "batch_ms": round(batch_s * 1e3, 1), "sequential_ms": round(seq_s * 1e3, 1),
           "speedup_vs_sequential": round(seq_s / batch_s, 2),
           "rhs_steps_per_s": round(NRHS * steps / batch_s, 2),
           "iterations_total": int(its.sum()), "iterations_max_per_step": [int(x) for x in its.max(axis=1)],
           "sequential_matches_batch_bitwise": bool(same), "setup_s": round(setup_s, 2)}
    if log:
        log(f"16 RHS x {steps} steps at N={n}: batched {batch_s * 1e3:.0f} ms, one after another {seq_s * 1e3:.0f} ms")
    op.cleanup()
    return out


def bench_solve_refined(n=320, steps=20, tol=1e-8, max_outer=12, max_inner=500, tol_inner=1e-4, log=None,
                        fp64_steps=None):
    """The sixteen right-hand sides as one 16-column refined solve: FP32 inner
    Jacobi-PCG on the matrix-core operator inside per-column FP64 defect
    correction, ABFT D1 on every inner apply.  The same steps of the fused
    FP64 batch give the reference the refined result's error is stated
    against."""
    from .heat import HeatBatch, HeatOperator
    k, rhoc = workload_fields(n)
    dt = 1e-4
    op = HeatOperator(n, dt, k.ravel(), rhoc.ravel(), kernel="fast", scatter="colored", source="const",
                      source_value=1.0, detect=True)
    F = load_vectors(n)
    batch = HeatBatch.shared(op, F)
    its, d1s, times = [], [], []
    for st in range(steps):
        t0 = time.perf_counter()
        it, d1 = batch.step_mrhs(1, max_outer, max_inner, tol, tol_inner)
        times.append(time.perf_counter() - t0)
        its.append(it[0])
        d1s.append(d1[0])
        if log:
            log(f"refined step {st + 1}/{steps}: {times[-1] * 1e3:.0f} ms, inner iterations max {int(it.max())} "
                f"min {int(it.min())}, largest D1 {float(d1.max()):.2e}")
    refined = [op.download(batch.vector(i, "T")) for i in range(NRHS)]
    batch.cleanup()
    # the same steps in plain FP64 (fused batch) for the error figure
    ref = HeatBatch.shared(op, F)
    f64_s = 0.0
    nref = steps if fp64_steps is None else fp64_steps
    for st in range(nref):
        t0 = time.perf_counter()
        ref.step(1, max_inner * max_outer, tol)
        f64_s += time.perf_counter() - t0
        if log:
            log(f"FP64 step {st + 1}/{nref} done ({f64_s * 1e3:.0f} ms so far)")
    errs = []
    if nref == steps:
        for i in range(NRHS):
            t64 = op.download(ref.vector(i, "T"))
            errs.append(float(np.linalg.norm(refined[i] - t64) / max(np.linalg.norm(t64), 1e-300)))
    ref.cleanup()
    its, d1s = np.array(its), np.array(d1s)
    total = float(sum(times))
    out = {"n": n, "rhs": NRHS, "steps": steps, "tol": tol, "tol_inner": tol_inner, "max_outer": max_outer,
           "max_inner": max_inner, "nodes": op.nn, "refined_ms": round(total * 1e3, 1),
           "rhs_steps_per_s": round(NRHS * steps / total, 2),
           "inner_iterations_total": int(its.sum()),
           "inner_iterations_max_per_step": [int(x) for x in its.max(axis=1)],
           "d1_max": float(d1s.max()), "d1_threshold_fp32": 4.76e-8,
           "fp64_ms": round(f64_s * 1e3, 1), "fp64_steps": nref}
    if errs:
        out["rel_err_vs_fp64_max"] = max(errs)
    if log:
        log(f"16-column refined: {steps} steps in {total * 1e3:.0f} ms; FP64 batch {f64_s * 1e3:.0f} ms")
    op.cleanup()
    return out


if __name__ == "__main__":
    main()
