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
    out = {
        "n": n, "nodes": nn, "elements": ne, "rhs": NRHS,
        "ms_per_apply": round(ms, 4),
        "gdof_rhs_per_s": round(nn * NRHS / (ms * 1e-3) / 1e9, 2),
        "algorithmic_bytes": alg_bytes,
        "achieved_GBs": round(alg_bytes / (ms * 1e-3) / 1e9, 1),
        "hbm_frac": round(alg_bytes / (ms * 1e-3) / 8e12, 4),
        "mfma_tflops": round(flops / (ms * 1e-3) / 1e12, 2),
        "mfma_frac_fp32": round(flops / (ms * 1e-3) / 157.3e12, 4),
        "single_fp32_ms_x16": round(16 * ms32, 3),
        "single_fp64_ms_x16": round(16 * ms64, 3),
        "speedup_vs_16_fp32_applies": round(16 * ms32 / ms, 2),
    }
    if log:
        log(f"16-RHS apply at N={n}: {ms:.3f} ms ({out['gdof_rhs_per_s']} GDOF-RHS/s, {out['mfma_tflops']} TF/s), "
            f"16 FP32 applies {16 * ms32:.3f} ms, 16 FP64 applies {16 * ms64:.3f} ms")
    op.cleanup()
    return out


def main():
    log = lambda m: print(f"[mrhs] {m}", file=sys.stderr, flush=True)
    print(json.dumps(bench_apply(log=log)))


if __name__ == "__main__":
    main()
