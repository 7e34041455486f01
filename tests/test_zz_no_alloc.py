"""heat_dot_static: the dot product of heat_dot with its block partials in a
file-scope device array instead of handle-owned memory, for callers that
cannot allocate.  Its result must equal heat_dot's bit for bit on the same
vectors (the same kernels, the same fixed order), call after call: the
array keeps the last call's contents, so a second, smaller call must not
read what the first left behind.
"""
import ctypes as C
import math

import numpy as np
import pytest

from mnhack8_amd import fields
from mnhack8_amd.heat import HeatOperator, load_library

MAX_N = 256 * 32768   # HEAT_STATIC_DOT_MAX_BLOCKS blocks of 256 values


def _dot_static(lib, op, x, y, n, out):
    rc = lib.heat_dot_static(op.ptr(x), op.ptr(y), n, out, op.stream())
    assert rc == 0, lib.heat_last_error()
    return op.download(out, 1)[0]


@pytest.mark.gpu
def test_dot_static_equals_heat_dot():
    lib = load_library()
    # the second mesh is smaller: its call uses fewer partial slots than the first left
    for n, seed in ((64, 1), (11, 2)):
        print(f"[case] heat_dot_static N={n} seed={seed}", flush=True)
        k, r = fields.lognormal(n, sigma=1.0, L=0.125, seed=seed)
        op = HeatOperator(n, 1e-3, k.ravel(), r.ravel())
        rng = np.random.default_rng(seed)
        x, y = rng.standard_normal(op.nn), rng.standard_normal(op.nn)
        op.upload("p", x)
        op.upload("z", y)
        ref = op.dot("p", "z")
        out = op.ptr("Ap")
        got = _dot_static(lib, op, "p", "z", op.nn, out)
        assert got == ref, (n, got, ref)
        assert abs(got - math.fsum(x * y)) <= 1e-12 * math.fsum(np.abs(x * y))
        # x.x through both paths
        assert _dot_static(lib, op, "p", "p", op.nn, out) == op.dot("p", "p")
        # a prefix of the vectors: the first 1000 values, then none
        assert abs(_dot_static(lib, op, "p", "z", 1000, out) - math.fsum(x[:1000] * y[:1000])) <= 1e-12 * 1000
        assert _dot_static(lib, op, "p", "z", 0, out) == 0.0
        # sizes beyond the static partials are refused before any work is queued
        for bad in (MAX_N + 1, -1):
            assert lib.heat_dot_static(op.ptr("p"), op.ptr("z"), bad, out, op.stream()) == 1
            assert b"outside" in lib.heat_last_error()
        op.cleanup()
