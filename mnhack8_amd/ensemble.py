"""Ensemble of independent heat solves, advanced together on one GPU.

The reference runs an ensemble as one solver process per member
(scripts/ensemble/run_ensemble.py:53-104, each a src/heat_solver_het.cu
main).  Here all members live in one HeatBatch: every PCG iteration of every
member is one pair of launches, members that have converged (or are masked)
return at once, and each member's iterates are bit-identical to its solve
alone.

Workload (the ensemble row of the measurement plan): COUNT members, mesh size
drawn uniformly from {16, 24, 32, 48, 64, 96}, kappa log-normal with sigma
drawn from {0.5, 1, 2} and the reference's correlated model (L = 0.125,
scripts/ensemble/run_ensemble.py:35-36), rho*c = 1, f = 1, T^0 = 0,
dt = 1e-3, 20 steps, PCG to 1e-8: FP32-refined (FP32 inner Jacobi-PCG to
1e-4 inside FP64 defect correction), or plain FP64 with --precision fp64.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

SIZES = (16, 24, 32, 48, 64, 96)
SIGMAS = (0.5, 1.0, 2.0)


def members(count=96, seed=1, dt=1e-3, sizes=SIZES, sigmas=SIGMAS):
    from . import fields
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        sigma = float(rng.choice(sigmas))
        k, rhoc = fields.lognormal(n, sigma=sigma, L=0.125, seed=1000 * seed + i)
        out.append({"n": n, "dt": dt, "k": k, "rhoc": rhoc, "source": "const", "source_value": 1.0,
                    "sigma": sigma, "seed": 1000 * seed + i})
    return out


def run(count=96, steps=20, tol=1e-8, max_iter=500, seed=1, sequential=True, log=None, precision="refined",
        tol_inner=1e-4, max_outer=50, persistent=True, autotune=True):
    """Time the batch (and, for comparison, the same members one after the
    other through the single-solve API) and return a summary.  precision
    "refined": every solve FP32-refined (iteration counts are inner ones);
    "fp64": plain FP64 PCG.  persistent: also time the same steps with each
    batched solve as one persistent (cooperative) launch instead of
    graph-replayed chunks, and check that T is the same bits.  autotune:
    first pick the batch's tile geometry (HeatBatch.autotune); the member
    handles, and so the one-after-another solves, run the same geometry."""
    refined = precision == "refined"
    if precision not in ("refined", "fp64"):
        raise ValueError("precision must be 'refined' or 'fp64'")

    def batch_steps(batch, nsteps):
        if refined:
            return batch.step_mixed(nsteps, max_outer, max_iter, tol, tol_inner)
        return batch.step(nsteps, max_iter, tol)

    def solve_one(op):
        if refined:
            # the batch's refined steps run the standard inner iteration
            op.set_inner_solver("standard")
            return op.pcg_solve_mixed("b", "T", max_outer, max_iter, tol, tol_inner)[1]
        return op.pcg_solve("b", "T", max_iter, tol)

    from .heat import HeatBatch
    ms = members(count, seed)
    t0 = time.perf_counter()
    batch = HeatBatch(ms)
    setup_s = time.perf_counter() - t0
    nodes = sum(op.nn for op in batch.ops)
    zero = [np.zeros(op.nn) for op in batch.ops]
    for op, z in zip(batch.ops, zero):
        op.upload("T", z)
        op.synchronize()
    schedule = {"planes": 16, "wave": 768, "tuned": False}
    if autotune:
        t0 = time.perf_counter()
        planes, wave, ms_it = batch.autotune("refined" if refined else "fp64", 20)
        schedule = {"planes": planes, "wave": wave, "tuned": True, "ms_per_iteration": round(ms_it, 4),
                    "tuning_s": round(time.perf_counter() - t0, 3),
                    "chunks_per_column": sorted({batch.geometry(i)[2] for i in range(count)})}
        if log:
            log(f"autotuned tile geometry: {planes} planes per chunk, wave fill {wave} "
                f"({ms_it * 1e3:.1f} us per batched iteration)")
    batch_steps(batch, 1)                        # warm-up (graph capture, clocks)
    for op, z in zip(batch.ops, zero):
        op.upload("T", z)
        op.synchronize()
    t0 = time.perf_counter()
    its = batch_steps(batch, steps)
    batch_s = time.perf_counter() - t0
    if log:
        log(f"batch ({precision}): {count} members, {nodes} nodes, {steps} steps in {batch_s * 1e3:.1f} ms")
    dof_it = int(sum(op.nn * int(its[:, i].sum()) for i, op in enumerate(batch.ops)))
    out = {
        "members": count, "steps": steps, "nodes": int(nodes), "tol": tol, "precision": precision,
        "dof_iterations": dof_it,
        "gdof_iterations_per_s": round(dof_it / batch_s / 1e9, 3),
        "sizes": sorted({m["n"] for m in ms}),
        "batch_ms": round(batch_s * 1e3, 3),
        "member_steps_per_s": round(count * steps / batch_s, 1),
        "dof_steps_per_s": round(nodes * steps / batch_s, 1),
        "iterations_total": int(its.sum()),
        "iterations_max_per_step": [int(x) for x in its.max(axis=1)],
        "setup_s": round(setup_s, 3),
        "tile_schedule": schedule,
    }
    if persistent:
        finals = [op.download("T") for op in batch.ops]
        batch.set_persistent(True)
        for op, z in zip(batch.ops, zero):
            op.upload("T", z)
            op.synchronize()
        batch_steps(batch, 1)                    # warm-up
        for op, z in zip(batch.ops, zero):
            op.upload("T", z)
            op.synchronize()
        t0 = time.perf_counter()
        pits = batch_steps(batch, steps)
        pers_s = time.perf_counter() - t0
        same = all(np.array_equal(op.download("T"), f) for op, f in zip(batch.ops, finals))
        batch.set_persistent(False)
        out.update({"persistent_ms": round(pers_s * 1e3, 3),
                    "persistent_matches_batch_bitwise": bool(same and np.array_equal(pits, its)),
                    "launch_form_kept": "persistent" if pers_s < batch_s else "graph-replayed chunks"})
        if log:
            log(f"persistent launch: {pers_s * 1e3:.1f} ms (graph-replayed chunks {batch_s * 1e3:.1f} ms)")
    if sequential:
        finals = [op.download("T") for op in batch.ops]
        t0 = time.perf_counter()
        seq_its = 0
        for op, z in zip(batch.ops, zero):
            op.upload("T", z)
            op.synchronize()
        t0 = time.perf_counter()
        for op in batch.ops:
            for _ in range(steps):
                op.build_rhs("T", "b")
                seq_its += solve_one(op)
            op.synchronize()
        seq_s = time.perf_counter() - t0
        same = all(np.array_equal(op.download("T"), f) for op, f in zip(batch.ops, finals))
        out.update({"sequential_ms": round(seq_s * 1e3, 3), "speedup_vs_sequential": round(seq_s / batch_s, 2),
                    "sequential_iterations_total": int(seq_its),
                    "sequential_matches_batch_bitwise": bool(same)})
        if log:
            log(f"sequential: {seq_s * 1e3:.1f} ms ({seq_s / batch_s:.2f}x the batch)")
    batch.cleanup()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--members", type=int, default=96)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--precision", choices=("refined", "fp64"), default="refined")
    ap.add_argument("--no-persistent", action="store_true", help="skip the persistent-launch timing")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile geometry")
    a = ap.parse_args(argv)
    log = lambda m: print(f"[ensemble] {m}", file=sys.stderr, flush=True)
    print(json.dumps(run(a.members, a.steps, a.tol, seed=a.seed, sequential=not a.no_sequential, log=log,
                         precision=a.precision, persistent=not a.no_persistent, autotune=not a.no_autotune)))


if __name__ == "__main__":
    main()
