#!/usr/bin/env python3
"""Fault-injection campaign over the heat solver: the grid, the runner and
the resumable ensemble.csv of the reference's scripts/ensemble/run_ensemble.py.

    python -m mnhack8_amd.campaign --out results/ensemble_N32 [--detect] [--mpi 1]
    python -m mnhack8_amd.aggregate --csv results/ensemble_N32/ensemble.csv --out results/ensemble_N32/agg

Same command line, grid and outputs as the reference:
* the grid (scripts/ensemble/run_ensemble.py:47-73): lognormal fields over
  --sigmas x --seeds; per field --targets random interior nodes (level G),
  elements (R) and face entries (H); per level the four bit classes with
  --reps bits drawn uniformly inside the class; injection at 10 %, 50 % and
  90 % of the run. All draws come from one random.Random(--rng) in the
  reference's order, so a grid and its run ids are the reference's.
* ensemble.csv (scripts/ensemble/run_ensemble.py:75-104): one row per run,
  the solver's fi_summary row plus run_id, field, sigma, seed, bit_class and
  status ("ok" or "error: ..."); runs already in the file are skipped, and
  --shard i/n keeps the rows with index % n == i.
* runs/<run_id>/ keeps fi_summary_e.csv, fi_trace_e.csv and meta_e.txt.

How the runs execute differs. The reference starts one solver process per
run, and each process repeats the two clean runs A and B of the field. Here
the pending runs of a field go to a few driver processes (heat_solver_het
--campaign), each of which loads the field once, runs A and B once and then
its share of the fault runs F; each F writes the same files as a single
invocation with --fi. A process advances up to --batch of its fault runs
together as one batch of solves (heat_batch_step: every PCG iteration of all
of them in one set of launches), and --jobs processes run at a time on the
GPU (small meshes leave it mostly idle). --per-run restores one process per
run (any binary with the reference's command line then works as --bin).

Level H (3) needs two ranks; this solver owns the whole mesh on one GPU, so
with --mpi >= 2 the level-H rows of the grid are recorded with an error
status, as a failed run is in the reference.
"""
from __future__ import annotations

import argparse
import csv
import itertools
import os
import random
import subprocess
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor, as_completed

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from mnhack8_amd import gen_field  # noqa: E402
from mnhack8_amd.validations import elem_index, node_index, read_csv  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mnhack8_amd", "_build", "heat_solver_het")

# bit classes of an FP64 word (scripts/ensemble/run_ensemble.py:24)
CLASSES = {"mant_lo": range(0, 30), "mant_hi": range(30, 52), "exp": range(52, 63), "sign": [63]}
ID_COLUMNS = {"run_id", "field", "sigma", "seed", "bit_class", "status"}


def injection_steps(t, dt):
    """Steps at 10, 50 and 90 % of the run (scripts/ensemble/run_ensemble.py:45)."""
    n_steps = int(round(t / dt))
    return sorted({max(1, int(f * n_steps)) for f in (0.1, 0.5, 0.9)})


def build_grid(N, t, dt, sigmas, seeds, reps, targets, levels, mpi, rng_seed, field_dir=""):
    """The run grid in the reference's order with the reference's random draws
    (scripts/ensemble/run_ensemble.py:44-70): per field, the node, element and
    face targets first, then per level and bit class REPS bit draws, each
    crossed with every injection step and target."""
    steps = injection_steps(t, dt)
    levels = [lv for lv in levels if not (lv == 3 and mpi < 2)]
    rng = random.Random(rng_seed)
    grid = []
    for sigma, seed in itertools.product(sigmas, seeds):
        tag = f"ln{N}_s{sigma}_seed{seed}"
        nodes = [node_index(rng.randint(2, N - 2), rng.randint(2, N - 2), rng.randint(2, N - 2), N)
                 for _ in range(targets)]
        elems = [elem_index(rng.randint(2, N - 3), rng.randint(2, N - 3), rng.randint(2, N - 3), N)
                 for _ in range(targets)]
        faces = [rng.randint(2, N - 2) * (N + 1) + rng.randint(2, N - 2) for _ in range(targets)]
        for level in levels:
            tg = {1: nodes, 2: elems, 3: faces}[level]
            for cls, bits in CLASSES.items():
                for _ in range(reps):
                    bit = rng.choice(list(bits))
                    for S in steps:
                        for ti, target in enumerate(tg):
                            grid.append(dict(field=tag, fp=os.path.join(field_dir, tag), sigma=sigma, seed=seed,
                                             level=level, bit_class=cls, bit=bit, step=S, target=target, tidx=ti))
    return grid


def run_id(g):
    return f"{g['field']}_L{g['level']}_{g['bit_class']}_b{g['bit']}_S{g['step']}_t{g['tidx']}"


def error_row(g, msg):
    """The row of a run that did not complete (scripts/ensemble/run_ensemble.py:94-96)."""
    return dict(run_id=run_id(g), field=g["field"], sigma=g["sigma"], seed=g["seed"], bit_class=g["bit_class"],
                fi_level=g["level"], fi_bit=g["bit"], fi_step=g["step"], status=f"error: {str(msg)[:120]}")


def ok_row(g, d):
    row = read_csv(os.path.join(d, "fi_summary_e.csv"))[0]
    row.update(run_id=run_id(g), field=g["field"], sigma=g["sigma"], seed=g["seed"], bit_class=g["bit_class"],
               status="ok")
    return row


class EnsembleCsv:
    """Append-only ensemble.csv. The header is the first row's columns
    (sorted); later rows are written under the file's header, so a row with
    fewer columns (an error row) leaves its other cells empty."""

    def __init__(self, path):
        self.path = path
        self.header = None
        if os.path.exists(path) and os.path.getsize(path) > 0:
            with open(path, newline="") as f:
                self.header = next(csv.reader(f))

    def done(self):
        return {r["run_id"] for r in read_csv(self.path)} if self.header else set()

    def append(self, row):
        if self.header is None:
            self.header = sorted(set(row.keys()) | ID_COLUMNS)
            with open(self.path, "w", newline="") as f:
                csv.writer(f).writerow(self.header)
        with open(self.path, "a", newline="") as f:
            csv.DictWriter(f, fieldnames=self.header, extrasaction="ignore").writerow(row)


def ensure_field(a, g):
    """The lognormal field of a grid row (scripts/ensemble/run_ensemble.py:82-83):
    --n N --model lognormal --sigma --L --upscale --seed, through --gen when given."""
    if os.path.exists(g["fp"] + "_k.bin"):
        return
    if a.gen:
        subprocess.run([sys.executable, a.gen, "--n", str(a.N), "--model", "lognormal", "--out", g["fp"],
                        "--sigma", str(g["sigma"]), "--L", str(a.L), "--upscale", str(a.upscale),
                        "--seed", str(g["seed"])], check=True, capture_output=True)
        return
    gen_field.main(["--n", str(a.N), "--model", "lognormal", "--out", g["fp"], "--sigma", str(g["sigma"]),
                    "--L", str(a.L), "--upscale", str(a.upscale), "--seed", str(g["seed"])])


def solver_args(a, fp, outdir):
    cmd = [a.bin, str(a.N), str(a.t), str(a.dt), fp, outdir, "--tag", "e", "--snap", "5",
           "--kernel", a.kernel, "--scatter", a.scatter]
    if a.detect:
        cmd.append("--detect")
    return cmd


_LOG_LOCK = threading.Lock()


def _log(path, cmd, r, dt):
    with _LOG_LOCK, open(path, "a") as f:
        f.write(f"$ {' '.join(cmd)}\n[{dt:.1f}s, rc={r.returncode}]\n{r.stdout}\n{r.stderr}\n")


def run_field(a, pending, log, part=0):
    """Pending runs of one field (distinct run ids) in one driver process;
    returns their rows in the given order. The field files must exist."""
    g0 = pending[0]
    spec = os.path.join(a.out, "runs", f"campaign_{g0['field']}_{part}.txt")
    rows = {}
    lines = []
    for g in pending:
        if g["level"] not in (1, 2):
            rows[run_id(g)] = error_row(g, "level H needs >= 2 ranks; one rank owns the whole mesh here")
            continue
        d = os.path.join(a.out, "runs", run_id(g))
        os.makedirs(d, exist_ok=True)
        summary = os.path.join(d, "fi_summary_e.csv")
        if os.path.exists(summary):
            os.remove(summary)
        lines.append(f"{d} {g['level']} {g['step']} {g['target']} {g['bit']}\n")
    err = ""
    if lines:
        with open(spec, "w") as f:
            f.writelines(lines)
        cmd = solver_args(a, g0["fp"], a.out) + ["--campaign", spec]
        if a.batch > 0:
            cmd += ["--campaign-batch", str(a.batch)]
        t0 = time.time()
        r = subprocess.run(cmd, capture_output=True, text=True)
        _log(log, cmd, r, time.time() - t0)
        if r.returncode != 0:
            err = f"command failed (rc={r.returncode}): {r.stderr[-200:]}"
    for g in pending:
        if run_id(g) in rows:
            continue
        d = os.path.join(a.out, "runs", run_id(g))
        try:
            rows[run_id(g)] = ok_row(g, d)
        except (OSError, IndexError, StopIteration) as ex:
            rows[run_id(g)] = error_row(g, err or ex)
    return [rows[run_id(g)] for g in pending]


def run_single(a, g, log):
    """One run as one solver process (scripts/ensemble/run_ensemble.py:84-96).
    The field files must exist."""
    d = os.path.join(a.out, "runs", run_id(g))
    os.makedirs(d, exist_ok=True)
    cmd = solver_args(a, g["fp"], d) + ["--fi", str(g["level"]), str(g["step"]), str(g["target"]), str(g["bit"])]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    _log(log, cmd, r, time.time() - t0)
    try:
        if r.returncode != 0:
            raise RuntimeError(f"command failed (rc={r.returncode}): {' '.join(cmd)}\n{r.stderr[-2000:]}")
        row = ok_row(g, d)
    except Exception as ex:  # noqa: BLE001 -- any failure becomes the run's error row
        row = error_row(g, ex)
    for fn in os.listdir(d):
        if fn.startswith("T_final_"):
            os.remove(os.path.join(d, fn))
    return row


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--bin", default=DRIVER)
    ap.add_argument("--gen", default="", help="field generator script (default: generate in-process)")
    ap.add_argument("--out", required=True)
    ap.add_argument("--N", type=int, default=32)
    ap.add_argument("--t", type=float, default=0.02)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--sigmas", default="0.5,1.0,2.0")
    ap.add_argument("--seeds", default="1,2,3")
    ap.add_argument("--L", type=float, default=0.125)
    ap.add_argument("--upscale", type=int, default=2)
    ap.add_argument("--reps", type=int, default=3, help="bit draws per class")
    ap.add_argument("--targets", type=int, default=3, help="interior targets per field")
    ap.add_argument("--levels", default="1,2,3")
    ap.add_argument("--mpi", type=int, default=2)
    ap.add_argument("--kernel", default="fast")
    ap.add_argument("--scatter", default="colored")
    ap.add_argument("--detect", action="store_true")
    ap.add_argument("--shard", default="0/1", help="i/n: run only rows with index %% n == i")
    ap.add_argument("--rng", type=int, default=12345)
    ap.add_argument("--per-run", action="store_true", help="one solver process per run, as the reference")
    ap.add_argument("--jobs", type=int, default=4,
                    help="solver processes at a time on the GPU (a field's runs are split between them)")
    ap.add_argument("--batch", type=int, default=128,
                    help="fault runs a driver process advances together as one batch (0: one after another)")
    return ap.parse_args(argv)


def main(argv=None):
    a = parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    field_dir = os.path.join(a.out, "fields")
    os.makedirs(field_dir, exist_ok=True)
    sigmas = [float(x) for x in a.sigmas.split(",")]
    seeds = [int(x) for x in a.seeds.split(",")]
    levels = [int(x) for x in a.levels.split(",")]
    grid = build_grid(a.N, a.t, a.dt, sigmas, seeds, a.reps, a.targets, levels, a.mpi, a.rng, field_dir)
    n_lv = len([lv for lv in levels if not (lv == 3 and a.mpi < 2)])
    print(f"grid: {len(grid)} runs ({len(sigmas)} sigmas x {len(seeds)} seeds x {n_lv} levels x "
          f"{len(CLASSES)} classes x {a.reps} reps x {len(injection_steps(a.t, a.dt))} steps x {a.targets} targets)",
          flush=True)
    shard_i, shard_n = (int(x) for x in a.shard.split("/"))
    out_csv = EnsembleCsv(os.path.join(a.out, "ensemble.csv"))
    done = out_csv.done()
    pending = [g for i, g in enumerate(grid) if i % shard_n == shard_i and run_id(g) not in done]
    log = os.path.join(a.out, "run.log")
    t0 = time.time()
    # every distinct run once (a grid can repeat a run id); rows go to the
    # csv in grid order when all runs are back
    distinct = list({run_id(g): g for g in pending}.values())
    for g in {g["field"]: g for g in distinct}.values():
        ensure_field(a, g)
    if a.per_run:
        tasks = [[g] for g in distinct]
        work = lambda part, gs: [run_single(a, gs[0], log)]
    else:
        fields_ = list(dict.fromkeys(g["field"] for g in distinct))
        per_field = max(1, -(-max(1, a.jobs) // max(1, len(fields_))))
        tasks = []
        for field in fields_:
            gs = [g for g in distinct if g["field"] == field]
            k = min(per_field, len(gs))
            tasks += [gs[i * len(gs) // k:(i + 1) * len(gs) // k] for i in range(k)]
        work = lambda part, gs: run_field(a, gs, log, part)
    rows, n_done = {}, 0
    with ThreadPoolExecutor(max_workers=max(1, a.jobs)) as pool:
        futures = {pool.submit(work, i, gs): gs for i, gs in enumerate(tasks)}
        for fut in as_completed(futures):
            gs = futures[fut]
            for g, row in zip(gs, fut.result()):
                rows[run_id(g)] = row
            n_done += len(gs)
            el = time.time() - t0
            print(f"[{n_done}/{len(distinct)} runs, {el / 60:.1f} min, {el / max(n_done, 1):.2f} s/run]", flush=True)
    for g in pending:
        out_csv.append(rows[run_id(g)])
    print(f"done: {len(pending)} new runs in {time.time() - t0:.1f} s -> {out_csv.path}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
