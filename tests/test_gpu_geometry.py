"""Tile geometry of the PCG kernels and its autotuners (heat_set_pcg_geometry,
heat_autotune_pcg, heat_batch_set_geometry, heat_batch_autotune) against the
CPU oracle.

The chunking of a tile column fixes the order of the per-tile inner-product
partials, so a solve's bits depend on the geometry; the oracle takes the same
(planes, wave) pair.  Bars: iteration counts exact, iterates, histories and
D1 records bit-identical to the oracle with the same geometry.
"""
import numpy as np
import pytest

from oracle_lib import OracleOp
from test_gpu_parity import BATCH_MEMBERS, _batch_and_oracles, heat, make_field

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (planes, wave): one plane pair per chunk, short and long chunks with and
# without the whole-wave fill (waves of 768 tiles, and of 1024 for the
# one-launch step's four blocks per CU), and one chunk for the whole column
GEOMETRIES = [(2, 0), (4, 0), (6, 768), (8, 1024), (32, 0), (64, 768), (1 << 20, 0)]


def _solve_both(op, orc, n, precision):
    b = orc.build_rhs(orc.initial_condition("sine", 0.0, 1.0))
    op.upload("b", b)
    op.upload("T", np.zeros(op.nn))
    op.abft_take()
    orc.abft_take()
    if precision == "fp64":
        got = op.pcg_solve("b", "T", 300, 1e-11, history=True)
        it, xo, histo = orc.pcg_fused(b, np.zeros(op.nn), 300, 1e-11)
        assert got[0] == it
        np.testing.assert_array_equal(got[1], histo)
    else:
        got = op.pcg_solve_mixed("b", "T", 50, 500, 1e-10, 1e-4)
        oo, io, xo = orc.pcg_mixed_cg1(b, np.zeros(op.nn), 50, 500, 1e-10, 1e-4)
        assert got == (oo, io)
    assert np.array_equal(op.download("T"), xo)
    d1, calls = op.abft_take()
    d1o, callso = orc.abft_take()
    assert (d1, calls) == (d1o, callso)
    return got


@pytest.mark.parametrize("n", [7, 33, 40])
@pytest.mark.parametrize("planes,wave", GEOMETRIES)
def test_pcg_geometry_bitwise(n, planes, wave):
    """The fused FP64 solve and the one-launch refined solve (D1 on) under a
    non-default geometry: bit-identical to the oracle given the same one."""
    dt = 1e-3
    print(f"[case] geometry n={n} planes={planes} wave={wave}", flush=True)
    k, rhoc = make_field(n, seed=70 + n)
    op = heat().HeatOperator(n, dt, k, rhoc, detect=True, source="const", source_value=1.0)
    orc = OracleOp(n, dt, k, rhoc, kernel="struct", detect=True, source="const", source_value=1.0)
    for which in (0, 1):
        orc.set_pcg_geometry(which, planes, wave)
    op.set_pcg_geometry("fused", planes, wave)
    op.set_pcg_geometry("one_launch", planes, wave)
    p, w, nch, tiles = op.pcg_geometry("fused")
    assert (p, w) == (planes, wave) and 1 <= nch <= (n + 2) // 2
    print(f"  fused chunks {nch} tiles {tiles}; one-launch {op.pcg_geometry('one_launch')}", flush=True)
    _solve_both(op, orc, n, "fp64")
    _solve_both(op, orc, n, "refined")
    # back to the default: the default bits again
    op.set_pcg_geometry("fused", 0)
    orc.set_pcg_geometry(0, 0, 0)
    _solve_both(op, orc, n, "fp64")
    op.cleanup()


def test_geometry_argument_checks():
    """Bad geometries are refused; a batch member's geometry is the batch's."""
    h = heat()
    n = 6
    k, rhoc = make_field(n, seed=5)
    op = h.HeatOperator(n, 1e-3, k, rhoc)
    for planes, wave in ((1, 0), (-4, 0), (8, -1)):
        with pytest.raises(h.HeatError):
            op.set_pcg_geometry("fused", planes, wave)
    assert op.pcg_geometry("fused")[:2] == (16, 768)
    op.cleanup()
    batch, _, _ = _batch_and_oracles(BATCH_MEMBERS[:2])
    with pytest.raises(h.HeatError):
        batch.ops[0].set_pcg_geometry("fused", 4, 0)
    with pytest.raises(h.HeatError):
        batch.ops[0].autotune_pcg("fused", 5)
    batch.set_geometry(4, 0)
    assert batch.geometry(0)[:2] == (4, 0) and batch.ops[0].pcg_geometry("fused")[:2] == (4, 0)
    batch.cleanup()


def test_autotune_pcg_keeps_bits_and_d1_record():
    """Both PCG autotuners pick a geometry, report it as the one in force,
    keep the handle's pending D1 record, and the solves that follow are
    bit-identical to the oracle's with the picked geometry."""
    n, dt = 40, 1e-3
    print("[case] autotune pcg", flush=True)
    k, rhoc = make_field(n, seed=91, sigma=1.5)
    op = heat().HeatOperator(n, dt, k, rhoc, detect=True, source="const", source_value=1.0)
    orc = OracleOp(n, dt, k, rhoc, kernel="struct", detect=True, source="const", source_value=1.0)
    x = np.random.default_rng(2).standard_normal(op.nn)
    op.upload("p", x)
    op.abft_take()
    orc.abft_take()
    op.apply_operator("p", "Ap", dt)            # a pending D1 record
    orc.apply(x, dt)
    picks = {}
    for which, code in (("fused", 0), ("one_launch", 1)):
        planes, wave, ms = op.autotune_pcg(which, 8)
        print(f"  {which}: planes {planes} wave {wave} ({ms * 1e3:.1f} us per iteration)", flush=True)
        assert op.pcg_geometry(which)[:2] == (planes, wave) and ms > 0
        orc.set_pcg_geometry(code, planes, wave)
        picks[which] = (planes, wave)
    assert op.abft_take() == orc.abft_take()
    _solve_both(op, orc, n, "fp64")
    _solve_both(op, orc, n, "refined")
    op.cleanup()


@pytest.mark.parametrize("mode", ["graph", "persistent"])
def test_batch_geometry_bitwise(mode):
    """An ensemble under a non-default geometry (FP64 steps, then refined
    steps under another): every member bit-identical to the oracle's solve
    of it alone with the same geometry."""
    print(f"[case] batch geometry {mode}", flush=True)
    batch, orcs, T0s = _batch_and_oracles(BATCH_MEMBERS)
    batch.set_tuning(4, True)
    batch.set_persistent(mode == "persistent")
    batch.set_geometry(4, 0)
    its = batch.step(2, 300, 1e-10)
    Ts = []
    for i, (op, orc, T0) in enumerate(zip(batch.ops, orcs, T0s)):
        orc.set_pcg_geometry(0, 4, 0)
        T = T0
        for s in range(2):
            it, T, _ = orc.pcg_fused(orc.build_rhs(T), T, 300, 1e-10)
            assert its[s, i] == it, (i, s)
        assert np.array_equal(op.download("T"), T)
        Ts.append(T)
    batch.set_geometry(32, 768)
    its = batch.step_mixed(1, 8, 200, 1e-10, 1e-4)
    for i, (op, orc, T) in enumerate(zip(batch.ops, orcs, Ts)):
        orc.set_pcg_geometry(0, 32, 768)
        _, inner, To = orc.pcg_mixed_fused(orc.build_rhs(T), T, 8, 200, 1e-10, 1e-4)
        assert its[0, i] == inner and np.array_equal(op.download("T"), To), i
    batch.cleanup()


def test_batch_autotune_restores_T():
    """The batch autotuner leaves every member's T as it found it (bit for
    bit), reports the geometry in force, and the steps that follow are
    bit-identical to the oracle's with that geometry."""
    print("[case] batch autotune", flush=True)
    batch, orcs, T0s = _batch_and_oracles(BATCH_MEMBERS)
    for mode in ("fp64", "refined"):
        planes, wave, ms = batch.autotune(mode, 6)
        print(f"  {mode}: planes {planes} wave {wave} ({ms * 1e3:.1f} us per iteration)", flush=True)
        for i, (op, T0) in enumerate(zip(batch.ops, T0s)):
            assert np.array_equal(op.download("T"), T0), i
            assert batch.geometry(i)[:2] == (planes, wave)
    its = batch.step(1, 300, 1e-10)
    for i, (op, orc, T0) in enumerate(zip(batch.ops, orcs, T0s)):
        orc.set_pcg_geometry(0, planes, wave)
        it, To, _ = orc.pcg_fused(orc.build_rhs(T0), T0, 300, 1e-10)
        assert its[0, i] == it and np.array_equal(op.download("T"), To), i
    batch.cleanup()
