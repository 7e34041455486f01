#!/usr/bin/env python3
"""Benchmark of the matrix-free Q1 heat solve on one MI355X.

Workload (the N=128 heterogeneous configuration): unit cube, 128^3 Q1 hex
elements (2,097,152 elements, 2,146,689 nodes), per-element conductivity
kappa = exp(clip(g, -3, 3)) with g ~ N(0,1) i.i.d. from a counter-based
generator keyed by (seed, element), rho*c = 1, f = 1, dt = 1e-4, FP64,
colour-ordered scatter.

A "step" is one operator apply y = P (M + dt kappa K) x (the PCG matvec),
timed on the device with HIP events over K back-to-back launches.  value =
nodes / time per apply (GDOF/s).  Also reported: the same with the caches
flushed before every apply, time-to-1e-8 of the first time step's solve for
plain FP64 PCG and for the FP32-refined solve, the roofline position of the
matvec kernel, and the CPU oracle's matvec rate on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, AMD specification
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "colored_matvec_n128_traffic.txt")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(schedule):
    """HBM bytes per launch of the matvec kernel of `schedule` ("tdma/p8/s0")
    from the committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, each
    scaled by the factor the same counter reads for a kernel of known bytes)."""
    try:
        with open(PROFILE_TRAFFIC) as f:
            for line in f:
                m = re.match(r"\s*traffic_bytes_per_launch\[([^\]]+)\]\s*[:=]\s*([0-9.eE+]+)", line)
                if m and m.group(1) == schedule:
                    return float(m.group(2))
    except OSError:
        pass
    return None


def cpu_baseline(n, k, rhoc, dt, reps=20, warm=3):
    """The CPU oracle's colour-sweep matvec (same arithmetic), OpenMP threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleOp
    # every core this process may run on, unless OMP_NUM_THREADS names fewer
    # (the GPU box sets it to the host share of one GPU)
    avail = len(os.sched_getaffinity(0))
    threads = max(1, min(avail, int(os.environ.get("OMP_NUM_THREADS", avail) or avail)))
    orc = OracleOp(n, dt, k, rhoc, kernel="struct", scatter="colored", source="const",
                   source_value=1.0, threads=threads)
    x = np.random.default_rng(0).standard_normal(orc.nn)
    for _ in range(warm):
        orc.apply(x, dt)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        orc.apply(x, dt)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    res = {"value": orc.nn / med / 1e9, "unit": "GDOF/s", "cores": threads, "kind": "port",
           "sample": f"{reps} colour-sweep matvecs of the same N={n} mesh after {warm} warm-ups, median "
                     f"{med * 1e3:.1f} ms per apply",
           "host_cpus": os.cpu_count(), "affinity_cpus": avail,
           "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    # the first time step's FP64 PCG to 1e-8 (T^0 = 0, b = dt F) on the same threads
    b = orc.build_rhs(np.zeros(orc.nn))
    t0 = time.perf_counter()
    its, _, _ = orc.pcg(b, np.zeros(orc.nn), 5000, 1e-8)
    res["time_to_1e-8"] = {"ms": round((time.perf_counter() - t0) * 1e3, 1), "iterations": its,
                           "cores": threads}
    # the matvec on one core
    orc.set_threads(1)
    orc.apply(x, dt)
    one = []
    for _ in range(3):
        t0 = time.perf_counter()
        orc.apply(x, dt)
        one.append(time.perf_counter() - t0)
    med1 = float(np.median(one))
    res["one_core"] = {"value": orc.nn / med1 / 1e9, "unit": "GDOF/s", "cores": 1,
                       "sample": f"3 colour-sweep matvecs after 1 warm-up, median {med1 * 1e3:.1f} ms per apply"}
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-solve", action="store_true", help="skip the time-to-tolerance solves")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--no-mrhs", action="store_true", help="skip the 16-RHS N=320 apply measurement")
    ap.add_argument("--no-hbm", action="store_true", help="skip the HBM-resident N=256 apply measurement")
    ap.add_argument("--seu-sites", type=int, default=2000,
                    help="sites per precision of the N=128 PCG-state upset campaign (0: skip it)")
    ap.add_argument("--sweep", action="store_true", help="time every apply schedule and print the table")
    ap.add_argument("--probe-copy", type=int, default=0,
                    help="run this many axpy passes (2 node vectors read, 1 written) first: the known-bytes "
                         "kernel that calibrates the PMC byte counters")
    ap.add_argument("--schedule", default=None,
                    help="impl/pPLANES/sSHAPE (e.g. tdma/p8/s0) instead of the autotuner's pick")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {a.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from mnhack8_amd import fields
    from mnhack8_amd.heat import HeatOperator

    n, dt = a.n, 1e-4
    log(f"generating fields N={n}")
    k, rhoc = fields.iid_lognormal(n, sigma=1.0, clip=3.0, seed=a.seed)
    k, rhoc = k.ravel(), rhoc.ravel()
    log("setup")
    op = HeatOperator(n, dt, k, rhoc, kernel="fast", scatter="colored", source="const", source_value=1.0)
    nn, ne = op.nn, op.ne
    x = np.random.default_rng(a.seed).standard_normal(nn)
    op.upload("p", x)

    for _ in range(a.probe_copy):
        op.axpy(0.5, "p", "r")
    op.synchronize()
    if a.sweep:
        table = {}
        cands = [("gather", 8, 0), ("march", 4, 0), ("march", 8, 0)] + \
                [(im, p, sh) for im in ("tile", "tdma") for sh in (0, 1, 2) for p in (4, 8, 16)]
        for impl, planes, shape in cands:
            op.set_colored_impl(impl)
            op.set_march_chunk(planes)
            op.set_tile_shape(shape)
            op.time_matvec("p", "Ap", 3)
            ms, _ = op.time_matvec("p", "Ap", 100)
            table[f"{impl}/p{planes}/s{shape}"] = round(ms * 1e3, 2)
        log(f"schedules (us/apply): {table}")
    if a.schedule:
        impl, pl, sh = a.schedule.split("/")
        op.set_colored_impl(impl)
        op.set_march_chunk(int(pl[1:]))
        op.set_tile_shape(int(sh[1:]))
        op.time_matvec("p", "Ap", 5)
        ms_pick, _ = op.time_matvec("p", "Ap", 50)
        from mnhack8_amd.heat import TILE_SHAPES
        pick = (impl, int(pl[1:]), TILE_SHAPES[int(sh[1:])], ms_pick)
    else:
        pick = op.autotune(50)
    picked = f"{pick[0]} planes={pick[1]} tile={pick[2][0]}x{pick[2][1]}"
    from mnhack8_amd.heat import TILE_SHAPES as _TS
    shape_id = {v: k for k, v in _TS.items()}.get(tuple(pick[2]), 0)
    sched_key = f"{pick[0]}/p{pick[1]}/s{shape_id}"
    log(f"autotuner picked {picked} ({pick[3] * 1e3:.2f} us)")
    log(f"warmup {a.warmup}")
    op.time_matvec("p", "Ap", max(1, a.warmup))
    torch.cuda.synchronize()
    log(f"timing {a.steps} applies")
    t_host0 = time.perf_counter()
    avg_ms, _ = op.time_matvec("p", "Ap", a.steps)
    torch.cuda.synchronize()
    host_s = time.perf_counter() - t_host0
    gdofs = nn / (avg_ms * 1e-3) / 1e9
    log(f"warm: {avg_ms * 1e3:.2f} us/apply  {gdofs:.1f} GDOF/s  (host wall {host_s * 1e3:.1f} ms)")

    cold_ms, cold_min = op.time_matvec("p", "Ap", 20, flush_bytes=1 << 30)
    log(f"cold (1 GiB flush before each): {cold_ms * 1e3:.2f} us/apply")

    alg_bytes = 8 * (2 * nn + 2 * ne)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    # the same bytes as a plain 16-byte stream (no stencil, no halo): how
    # close the apply comes to what this memory system streams at this size
    op.time_stream("p", "Ap", 20)
    stream_ms = op.time_stream("p", "Ap", 500)
    stream_gbs = alg_bytes / (stream_ms * 1e-3) / 1e9
    log(f"stream of the same bytes: {stream_ms * 1e3:.2f} us ({stream_gbs:.0f} GB/s); "
        f"apply at {stream_ms / avg_ms:.3f} of it")
    achieved_cold = alg_bytes / (cold_ms * 1e-3) / 1e9
    traffic = read_traffic(sched_key)

    result = {
        "metric": "PCG matvec GDOF/s at N=128 hex; time-to-1e-8, FP32-refined vs FP64",
        "value": round(gdofs, 3),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": avg_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"N={n} Q1 hex unit cube, iid log-normal kappa clip 3, rho*c=1, f=1, dt=1e-4, "
                               f"FP64 colour-ordered matvec y = P(M + dt kappa K)x",
                   "n": n, "nodes": nn, "elements": ne, "scatter": "colored (node-centric, colour order)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"k_apply_tdma ({sched_key})", "us_per_launch": round(avg_ms * 1e3, 2),
                     "timing": "mean of the timed back-to-back launches, HIP events on the library's stream: "
                               "the per-dispatch duration of a kernel trace plus the 0.1-0.2 us between "
                               "dispatches (profiles/colored_matvec_n128_dispatch.txt)",
                     "cache_resident": "the N=128 working set (68 MB per apply) stays in the 256 MB Infinity "
                                       "Cache; cold_cache and hbm_resident give the HBM-bound figures"},
        "cold_cache": {"ms_per_step": cold_ms, "value": round(nn / (cold_ms * 1e-3) / 1e9, 3),
                       "achieved_GBs": round(achieved_cold, 1), "frac": round(achieved_cold / HBM_PEAK_GBS, 4)},
        "stream_ceiling": {"kernel": "k_stream_ceiling (x, rho*c*m3, dt*kappa*k2 read; y written; 16-byte units)",
                           "us_per_launch": round(stream_ms * 1e3, 2), "achieved_GBs": round(stream_gbs, 1),
                           "frac": round(stream_gbs / HBM_PEAK_GBS, 4),
                           "apply_vs_stream": round(stream_ms / avg_ms, 4)},
        "algorithmic_bytes_per_apply": alg_bytes,
        "schedule": {"picked": picked, "key": sched_key, "us_per_apply": round(pick[3] * 1e3, 2)},
    }

    # the reference's two element-centric scatter forms on the same mesh and
    # vector: the 8-launch colour sweep and the atomic scatter (both kept)
    op.set_colored_impl("sweep")
    op.time_matvec("p", "Ap", 3)
    ms_sweep, _ = op.time_matvec("p", "Ap", 50)
    y_col = op.download("Ap")
    op.set_colored_impl(pick[0])
    op.set_march_chunk(int(pick[1]))
    op.set_tile_shape(shape_id)
    opa = HeatOperator(n, dt, k, rhoc, kernel="fast", scatter="atomic", source="const", source_value=1.0)
    opa.upload("p", x)
    atomic = {}
    for impl in ("tiled", "element"):
        opa.set_atomic_impl(impl)
        opa.time_matvec("p", "Ap", 3)
        ms_a, _ = opa.time_matvec("p", "Ap", 100)
        y_at = opa.download("Ap")
        atomic[impl] = {"us_per_apply": round(ms_a * 1e3, 2), "gdof_s": round(nn / (ms_a * 1e-3) / 1e9, 2),
                        "frac": round(alg_bytes / (ms_a * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "max_rel_diff_vs_colour": float(np.max(np.abs(y_at - y_col)) / np.max(np.abs(y_col)))}
    opa.cleanup()
    result["scatter_compare"] = {
        "node_centric_colour_order": {"us_per_apply": round(avg_ms * 1e3, 2), "gdof_s": round(gdofs, 2)},
        "colour_sweep_8_launches": {"us_per_apply": round(ms_sweep * 1e3, 2),
                                    "gdof_s": round(nn / (ms_sweep * 1e-3) / 1e9, 2)},
        "atomic_scatter": dict(atomic["tiled"], kernel="k_apply_scatter (zero fill + element tiles summed in LDS, "
                                                       "global atomics where tiles meet)"),
        "atomic_scatter_element_kernel": dict(atomic["element"], kernel="k_apply_elem (one thread per element, "
                                                                        "8 global atomic adds each)"),
    }
    log(f"scatter forms: colour sweep {ms_sweep * 1e3:.1f} us, atomic tiled {atomic['tiled']['us_per_apply']} us, "
        f"atomic element kernel {atomic['element']['us_per_apply']} us")

    if not a.no_hbm:
        # the same apply where its data cannot stay in the Infinity Cache
        # (N=256: 0.54 GB per apply, twice the cache), beside the stream of
        # the same bytes at that size
        n2 = 256
        log(f"HBM-resident apply at N={n2}")
        k2_, r2_ = fields.iid_lognormal(n2, sigma=1.0, clip=3.0, seed=a.seed)
        op2 = HeatOperator(n2, dt, k2_.ravel(), r2_.ravel(), kernel="fast", scatter="colored", source="const",
                           source_value=1.0)
        del k2_, r2_
        op2.upload("p", np.random.default_rng(a.seed).standard_normal(op2.nn))
        pick2 = op2.autotune(20)
        op2.time_matvec("p", "Ap", 5)
        disp2, _ = op2.time_matvec("p", "Ap", 100)
        op2.time_stream("p", "Ap", 10)
        st2 = op2.time_stream("p", "Ap", 100)
        bytes2 = 8 * (2 * op2.nn + 2 * op2.ne)
        result["hbm_resident"] = {
            "n": n2, "nodes": op2.nn, "elements": op2.ne, "algorithmic_bytes_per_apply": bytes2,
            "schedule": f"{pick2[0]} planes={pick2[1]} tile={pick2[2][0]}x{pick2[2][1]}",
            "us_per_launch": round(disp2 * 1e3, 2), "gdof_s": round(op2.nn / (disp2 * 1e-3) / 1e9, 2),
            "frac": round(bytes2 / (disp2 * 1e-3) / 8e12, 4),
            "stream_us": round(st2 * 1e3, 2), "stream_frac": round(bytes2 / (st2 * 1e-3) / 8e12, 4),
            "apply_vs_stream": round(st2 / disp2, 4)}
        log(f"N={n2}: {disp2 * 1e3:.1f} us per launch (frac {result['hbm_resident']['frac']}), stream "
            f"{st2 * 1e3:.1f} us (frac {result['hbm_resident']['stream_frac']})")
        op2.cleanup()

    if not a.no_solve:
        # time to relative residual 1e-8 for the first time step (T^0 = 0, b = dt F)
        op.upload("T", np.zeros(nn))
        op.build_rhs("T", "b")
        op.synchronize()
        # reference solution to a tight tolerance for the error figures
        op.pcg_solve("b", "T", 5000, 1e-13)
        xref = op.download("T")
        res = {}
        # each solve on its own autotuned tile geometry (chunking of the PCG tiles)
        geo = {}
        for which in ("fused", "one_launch"):
            planes, wave, ms_it = op.autotune_pcg(which, 20)
            geo[which] = {"planes": planes, "wave": wave, "chunks_per_column": op.pcg_geometry(which)[2],
                          "us_per_iteration": round(ms_it * 1e3, 2)}
            log(f"PCG geometry ({which}): {planes} planes per chunk, wave fill {wave} "
                f"({ms_it * 1e3:.1f} us per iteration)")
        res["tile_geometry"] = geo
        runs = (("fp64", "fused"), ("fp32_refined", "fused"),
                ("fp64_reference_order", "reference"), ("fp32_refined_reference_order", "reference"))
        for label, mode in runs:
            op.set_pcg_mode(mode)
            best = None
            for rep in range(7):
                op.upload("T", np.zeros(nn))
                op.synchronize()
                t0 = time.perf_counter()
                if label.startswith("fp64"):
                    its = op.pcg_solve("b", "T", 5000, 1e-8)
                    extra = {"iterations": its}
                else:
                    outer, inner = op.pcg_solve_mixed("b", "T", 50, 5000, 1e-8, 1e-4)
                    extra = {"outer": outer, "inner_iterations": inner}
                op.synchronize()
                ms = (time.perf_counter() - t0) * 1e3
                if best is None or ms < best:
                    best = ms
            xs = op.download("T")
            err = float(np.linalg.norm(xs - xref) / np.linalg.norm(xref))
            res[label] = {"ms": round(best, 3), "rel_err_vs_fp64_1e-13": err, **extra}
            log(f"time-to-1e-8 {label}: {best:.2f} ms {extra} err {err:.2e}")
        op.set_pcg_mode("fused")
        res["speedup_fp32_refined"] = round(res["fp64"]["ms"] / res["fp32_refined"]["ms"], 3)
        res["speedup_fused_vs_reference_order"] = round(res["fp64_reference_order"]["ms"] / res["fp64"]["ms"], 3)
        result["time_to_1e-8"] = res

    if a.seu_sites > 0 and not a.no_solve:
        # the upset campaign of the mixed-precision configuration: the same
        # sites for the FP64 and the refined solve, plus the refined solve's
        # own FP64 state, coverage with Wilson intervals
        from mnhack8_amd import seu_campaign
        log(f"SEU campaign: {a.seu_sites} sites per precision at N={n}")
        t0 = time.perf_counter()
        camp = seu_campaign.Campaign(n=n, seed=a.seed)
        clean = camp.clean()
        rows = camp.run(a.seu_sites, seed=a.seed, log=log)
        summ = seu_campaign.summarize(rows, clean)
        keep = ("sites", "benign", "detected", "silent", "coverage", "coverage_ci95", "coverage_D1_only",
                "by_detector")
        result["seu_campaign"] = {
            "sites_per_precision": a.seu_sites, "wall_s": round(time.perf_counter() - t0, 1),
            "thresholds": clean["thresholds"],
            "fp64": {k: summ["fp64"][k] for k in keep},
            "fp32_refined": {k: summ["fp32_refined"][k] for k in keep},
            "fp32_refined_whole_state": {k: summ["fp32_refined_whole_state"][k] for k in keep},
            "coverage_difference_pp": summ["coverage_difference_pp"],
            "coverage_difference_ci95_pp": summ["coverage_difference_ci95_pp"],
            "d1_coverage_difference_pp": summ["d1_coverage_difference_pp"],
            "d1_coverage_difference_ci95_pp": summ["d1_coverage_difference_ci95_pp"]}
        log(f"SEU campaign: coverage FP64 {summ['fp64']['coverage']:.3f}, refined (whole state) "
            f"{summ['fp32_refined_whole_state']['coverage']:.3f}, difference "
            f"{summ['coverage_difference_pp']:+.1f} pp")
        camp.op.cleanup()

    if not a.no_ensemble:
        from mnhack8_amd import ensemble
        log("ensemble: 96 members x 20 steps, FP32-refined, batched and one after the other")
        result["ensemble"] = ensemble.run(96, 20, 1e-8, log=log, precision="refined")
        log("ensemble: the same with plain FP64 PCG")
        result["ensemble_fp64"] = ensemble.run(96, 20, 1e-8, log=log, precision="fp64")

    if not a.no_mrhs:
        from mnhack8_amd import mrhs
        log("16-RHS apply at N=320 (matrix-core contraction)")
        result["multi_rhs"] = {"apply": mrhs.bench_apply(320, log=log),
                               "solve": mrhs.bench_solve(320, steps=1, log=log)}
        log("16-column refined solve at N=320 (matrix-core operator, FP64 defect correction, D1 on): 1 step")
        result["multi_rhs"]["solve_refined"] = mrhs.bench_solve_refined(320, steps=1, log=log)

    if not a.no_cpu:
        try:
            log("CPU baseline")
            result["cpu_baseline"] = cpu_baseline(n, k, rhoc, dt)
        except Exception as ex:  # the oracle is optional on a user's machine
            result["cpu_baseline"] = {"value": None, "error": str(ex)}
    op.cleanup()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
