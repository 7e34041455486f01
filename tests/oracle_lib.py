"""ctypes access to the CPU oracle (oracle/_build/libheat_oracle.so) for the
tests.  The oracle is the checker, never the thing measured."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libheat_oracle.so")
ORACLE_CLI = os.path.join(ROOT, "oracle", "_build", "heat_oracle")

KERNEL = {"gauss": 0, "fast": 1, "struct": 2}
_P, _D, _I, _L = C.c_void_p, C.c_double, C.c_int, C.c_long
_SIG = {
    "ho_create": (_P, [_I, _D, _P, _P, _I, _I, _I, _I, _D, _I]),
    "ho_destroy": (None, [_P]),
    "ho_last_error": (C.c_char_p, []),
    "ho_set_threads": (None, [_P, _I]),
    "ho_set_no_source": (None, [_P, _I]),
    "ho_nn": (_L, [_P]),
    "ho_ne": (_L, [_P]),
    "ho_get_vector": (_I, [_P, _I, _P]),
    "ho_get_ref_matrices": (None, [_P, _P, _P]),
    "ho_get_stats": (None, [_P, _P]),
    "ho_apply_operator": (None, [_P, _P, _P, _D, _I, _I]),
    "ho_build_rhs": (None, [_P, _P, _P]),
    "ho_dot": (_D, [_P, _P, _P]),
    "ho_minmax": (None, [_P, _P, _P, _P]),
    "ho_apply_preconditioner": (None, [_P, _P, _P]),
    "ho_apply_operator_f32": (None, [_P, _P, _P, _D, _I]),
    "ho_arm_fault": (None, [_P, _I, _L, _I]),
    "ho_flip_bit": (None, [_P, _L, _I]),
    "ho_abft_take": (None, [_P, _P, _P]),
    "ho_pcg_solve": (_I, [_P, _P, _P, _I, _D, _P]),
    "ho_pcg_solve_fused": (_I, [_P, _P, _P, _I, _D, _P]),
    "ho_pcg_solve_mixed_fused": (_I, [_P, _P, _P, _I, _I, _D, _D, _P, _P]),
    "ho_pcg_solve_mixed_cg1": (_I, [_P, _P, _P, _I, _I, _D, _D, _P, _P]),
    "ho_fused_sb_dot": (_D, [_P, _P, _P]),
    "ho_set_vector": (_I, [_P, _I, _P]),
    "ho_arm_pcg_fault": (None, [_P, _I, _L, _I, _I]),
    "ho_pcg_fault_fired": (_I, [_P]),
    "ho_fused_tile_sums": (None, [_P, _P, _P, _P, _P]),
    "ho_initial_condition": (None, [_P, _P, _I, _D, _D]),
    "ho_proto_run": (_I, [_I, _D, _D, _I, _I, _P, _P]),
    "ho_set_pcg_geometry": (_I, [_P, _I, _I, _I]),
    "ho_mixed_status": (_I, [_P]),
}
_lib = None


def lib():
    global _lib
    if _lib is None:
        l = C.CDLL(ORACLE_SO)
        for k, (r, a) in _SIG.items():
            f = getattr(l, k)
            f.restype = r
            f.argtypes = a
        _lib = l
    return _lib


def _p(a):
    return a.ctypes.data


class OracleOp:
    def __init__(self, n, dt, k, rhoc, kernel="struct", scatter="colored", verify=False,
                 source="gaussian", source_value=0.0, detect=False, no_source=False, threads=1):
        self.k = np.ascontiguousarray(k, dtype=np.float64).ravel()
        self.rhoc = np.ascontiguousarray(rhoc, dtype=np.float64).ravel()
        h = lib().ho_create(n, dt, _p(self.k), _p(self.rhoc), KERNEL[kernel], int(scatter == "colored"),
                            int(verify), {"gaussian": 0, "const": 1}[source], float(source_value), int(detect))
        if not h:
            raise RuntimeError(lib().ho_last_error().decode())
        self.h = h
        self.n, self.dt = n, dt
        self.nn = lib().ho_nn(h)
        self.ne = lib().ho_ne(h)
        if no_source:
            lib().ho_set_no_source(h, 1)
        if threads > 1:
            lib().ho_set_threads(h, threads)

    def set_threads(self, threads):
        lib().ho_set_threads(self.h, int(threads))

    def set_pcg_geometry(self, which, planes, wave):
        """The GPU library's tile geometry (0 fused, 1 one-launch inner)."""
        if lib().ho_set_pcg_geometry(self.h, int(which), int(planes), int(wave)):
            raise ValueError("bad geometry")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ho_destroy(self.h)
            self.h = None

    def set_vector(self, which, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == self.nn
        if lib().ho_set_vector(self.h, {"diag": 0, "F": 1}[which], _p(v)):
            raise KeyError(which)

    def vector(self, which):
        out = np.empty(self.nn)
        if lib().ho_get_vector(self.h, {"diag": 0, "F": 1, "wM": 2, "wA": 3, "ones": 4}[which], _p(out)):
            raise KeyError(which)
        return out

    def apply(self, x, dt_eff, allow_fi=False, project=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.nn)
        lib().ho_apply_operator(self.h, _p(x), _p(y), float(dt_eff), int(allow_fi), int(project))
        return y

    def apply_f32(self, x, dt_eff, project=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(self.nn, dtype=np.float32)
        lib().ho_apply_operator_f32(self.h, _p(x), _p(y), float(dt_eff), int(project))
        return y

    def build_rhs(self, T):
        T = np.ascontiguousarray(T, dtype=np.float64)
        b = np.empty(self.nn)
        lib().ho_build_rhs(self.h, _p(T), _p(b))
        return b

    def dot(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        return lib().ho_dot(self.h, _p(x), _p(y))

    def minmax(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        a, b = C.c_double(), C.c_double()
        lib().ho_minmax(self.h, _p(x), C.byref(a), C.byref(b))
        return a.value, b.value

    def precondition(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty(self.nn)
        lib().ho_apply_preconditioner(self.h, _p(r), _p(z))
        return z

    def arm_fault(self, level, target, bit):
        lib().ho_arm_fault(self.h, level, target, bit)

    def abft_take(self):
        m, c = C.c_double(), C.c_long()
        lib().ho_abft_take(self.h, C.byref(m), C.byref(c))
        return m.value, c.value

    def pcg(self, b, x, max_iter=500, tol=1e-8):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        hist = np.full(3 * (max_iter + 1), np.nan)
        it = lib().ho_pcg_solve(self.h, _p(b), _p(x), max_iter, tol, _p(hist))
        return it, x, hist.reshape(-1, 3)[: it + 1]

    def pcg_fused(self, b, x, max_iter=500, tol=1e-8):
        """The same solve with the GPU library's fused summation order."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        hist = np.full(3 * (max_iter + 1), np.nan)
        it = lib().ho_pcg_solve_fused(self.h, _p(b), _p(x), max_iter, tol, _p(hist))
        return it, x, hist.reshape(-1, 3)[: it + 1]

    def pcg_mixed_fused(self, b, x, max_outer=50, max_inner=500, tol=1e-8, tol_inner=1e-4):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        o, i = C.c_int(), C.c_int()
        lib().ho_pcg_solve_mixed_fused(self.h, _p(b), _p(x), max_outer, max_inner, tol, tol_inner,
                                       C.byref(o), C.byref(i))
        return o.value, i.value, x

    def pcg_mixed_cg1(self, b, x, max_outer=50, max_inner=500, tol=1e-8, tol_inner=1e-4):
        """The refined solve with the GPU library's one-launch inner iteration."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        o, i = C.c_int(), C.c_int()
        lib().ho_pcg_solve_mixed_cg1(self.h, _p(b), _p(x), max_outer, max_inner, tol, tol_inner,
                                     C.byref(o), C.byref(i))
        return o.value, i.value, x

    def arm_pcg_fault(self, vec, node, bit, iteration):
        lib().ho_arm_pcg_fault(self.h, {"u": 0, "r": 1, "p": 2, "z": 3, "Ap": 4, "x64": 5, "Ax64": 6}[vec],
                               int(node), int(bit), int(iteration))

    def pcg_fault_fired(self):
        return bool(lib().ho_pcg_fault_fired(self.h))

    def mixed_status(self):
        """1 tolerance met, 2 outer residual not finite, 0 pass limit (last refined solve)."""
        return lib().ho_mixed_status(self.h)

    def fused_sb_dot(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        return lib().ho_fused_sb_dot(self.h, _p(x), _p(y))

    def fused_tile_sums(self, p, Ap, w):
        p, Ap, w = (np.ascontiguousarray(v, dtype=np.float64) for v in (p, Ap, w))
        out = np.empty(4)
        lib().ho_fused_tile_sums(self.h, _p(p), _p(Ap), _p(w), _p(out))
        return out

    def initial_condition(self, mode, t=0.0, alpha=1.0):
        v = np.empty(self.nn)
        lib().ho_initial_condition(self.h, _p(v), {"zero": 0, "gaussian": 1, "sine": 2}[mode], t, alpha)
        return v

    def ref_matrices(self):
        M = np.empty(64)
        K = np.empty(64)
        lib().ho_get_ref_matrices(self.h, _p(M), _p(K))
        return M.reshape(8, 8), K.reshape(8, 8)

    def stats(self):
        o = np.empty(4)
        lib().ho_get_stats(self.h, _p(o))
        return o


def proto_run(n, t_final, dt, threads=8):
    """The MNHack7 prototype run (quadrant material at each quadrature point):
    per-step CG iterations and ||T||_2."""
    cap = int(t_final / dt + 0.5) + 1
    its = np.zeros(cap, dtype=np.int32)
    nrm = np.zeros(cap)
    k = lib().ho_proto_run(n, t_final, dt, threads, cap, _p(its), _p(nrm))
    assert k > 0
    return its[:k], nrm[:k]


def flip_bit(v, idx, bit):
    v = np.array(v, dtype=np.float64, copy=True)
    u = v.view(np.uint64)
    u[idx] ^= np.uint64(1) << np.uint64(bit)
    return v
