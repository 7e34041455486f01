"""Python binding of the MI355X heat solver's C ABI (include/heat.h).

``HeatOperator`` mirrors the reference's operator class
(src/heat_solver_het.cu:618-989: setup, apply_operator, apply_matvec,
apply_preconditioner, build_rhs, dot, norm2, minmax, setup_detectors,
cleanup) and ``pcg_solve`` its solver (src/heat_solver_het.cu:994-1039).
Vector arguments are device pointers: an int, an object with ``data_ptr()``
(a torch CUDA tensor) or the name of a handle-owned vector ("T", "b", "r",
"z", "p", "Ap", "diag", "F", "wM", "wA", "ones").

The library is loaded from mnhack8_amd/_build/libheat_hip.so; there is no CPU
fallback: a missing library raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libheat_hip.so")
DRIVER_PATH = os.path.join(_HERE, "_build", "heat_solver_het")

KERNEL = {"gauss": 0, "fast": 1}
SCATTER = {"atomic": 0, "colored": 1}
SOURCE = {"gaussian": 0, "const": 1}
COLORED_IMPL = {"gather": 0, "sweep": 1, "march": 2, "tile": 3, "tdma": 4}
ATOMIC_IMPL = {"element": 0, "tiled": 1}
TILE_SHAPES = {0: (8, 32), 1: (4, 64), 2: (16, 16)}
SEU_VEC = {"u": 0, "r": 1, "p": 2, "z": 3, "Ap": 4, "x64": 5, "Ax64": 6}
VEC = {"T": 0, "b": 1, "r": 2, "z": 3, "p": 4, "Ap": 5, "diag": 6, "F": 7, "wM": 8, "wA": 9, "ones": 10}
ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "bad field", 4: "self-check failed", 5: "bad state"}


class HeatError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{ERRORS.get(code, code)}] {msg}")
        self.code = code


class HeatConfig(C.Structure):
    _fields_ = [("n", C.c_int), ("dt", C.c_double), ("kernel", C.c_int), ("scatter", C.c_int),
                ("verify", C.c_int), ("no_source", C.c_int), ("source_mode", C.c_int),
                ("source_value", C.c_double), ("detect", C.c_int)]


_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_I64 = C.c_int64
SIGNATURES = {
    "heat_last_error": (C.c_char_p, []),
    "heat_setup": (_I, [C.POINTER(HeatConfig), _P, _P, C.POINTER(_P)]),
    "heat_cleanup": (_I, [_P]),
    "heat_sizes": (_I, [_P, C.POINTER(_I64), C.POINTER(_I64)]),
    "heat_field_stats": (_I, [_P, C.POINTER(_D)]),
    "heat_vector": (_I, [_P, _I, C.POINTER(_P)]),
    "heat_stream": (_I, [_P, C.POINTER(_P)]),
    "heat_synchronize": (_I, [_P]),
    "heat_upload": (_I, [_P, _P, _P, _I64]),
    "heat_download": (_I, [_P, _P, _P, _I64]),
    "heat_set_colored_impl": (_I, [_P, _I]),
    "heat_set_atomic_impl": (_I, [_P, _I]),
    "heat_apply_operator": (_I, [_P, _P, _P, _D, _I, _I]),
    "heat_apply_operator_f32": (_I, [_P, _P, _P, _D, _I]),
    "heat_apply_matvec": (_I, [_P, _P, _P]),
    "heat_apply_preconditioner": (_I, [_P, _P, _P]),
    "heat_build_rhs": (_I, [_P, _P, _P]),
    "heat_dot": (_I, [_P, _P, _P, C.POINTER(_D)]),
    "heat_norm2": (_I, [_P, _P, C.POINTER(_D)]),
    "heat_minmax": (_I, [_P, _P, C.POINTER(_D), C.POINTER(_D)]),
    "heat_set_dirichlet": (_I, [_P, _P, _D]),
    "heat_copy": (_I, [_P, _P, _P]),
    "heat_axpy": (_I, [_P, _D, _P, _P]),
    "heat_initial_condition": (_I, [_P, _P, _I, _D, _D]),
    "heat_pcg_solve": (_I, [_P, _P, _P, _I, _D, C.POINTER(_I), _P]),
    "heat_pcg_solve_mixed": (_I, [_P, _P, _P, _I, _I, _D, _D, C.POINTER(_I), C.POINTER(_I)]),
    "heat_arm_fault": (_I, [_P, _I, _I64, _I]),
    "heat_disarm_fault": (_I, [_P]),
    "heat_flip_bit": (_I, [_P, _P, _I64, _I]),
    "heat_abft_take": (_I, [_P, C.POINTER(_D), C.POINTER(_I64)]),
    "heat_abft_threshold": (_I, [_P, C.POINTER(_D)]),
    "heat_set_tuning": (_I, [_P, _I, _I]),
    "heat_set_march_chunk": (_I, [_P, _I]),
    "heat_set_pcg_mode": (_I, [_P, _I]),
    "heat_arm_pcg_fault": (_I, [_P, _I, C.c_int64, _I, _I]),
    "heat_apply_mrhs": (_I, [_P, _P, _P, _D]),
    "heat_time_mrhs": (_I, [_P, _P, _P, _I, C.POINTER(_D)]),
    "heat_batch_setup": (_I, [_I, _P, _P, _P, C.POINTER(_P)]),
    "heat_batch_cleanup": (_I, [_P]),
    "heat_batch_setup_rhs": (_I, [_P, _I, _P, C.POINTER(_P)]),
    "heat_batch_vector": (_I, [_P, _I, _I, C.POINTER(_P)]),
    "heat_batch_instance": (_I, [_P, _I, C.POINTER(_P)]),
    "heat_batch_set_active": (_I, [_P, _P]),
    "heat_batch_set_tuning": (_I, [_P, _I, _I]),
    "heat_batch_set_persistent": (_I, [_P, _I]),
    "heat_batch_step": (_I, [_P, _I, _I, _D, _P]),
    "heat_batch_detect": (_I, [_P, _P, _P, _P, _P]),
    "heat_batch_step_mixed": (_I, [_P, _I, _I, _I, _D, _D, _P]),
    "heat_batch_step_mrhs": (_I, [_P, _I, _I, _I, _D, _D, _P, _P]),
    "heat_pcg_fault_fired": (_I, [_P, C.POINTER(_I)]),
    "heat_pcg_mixed_status": (_I, [_P, C.POINTER(_I)]),
    "heat_autotune": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_D)]),
    "heat_set_tile_shape": (_I, [_P, _I]),
    "heat_time_matvec": (_I, [_P, _P, _P, _I, _I64, C.POINTER(_D), C.POINTER(_D)]),
    "heat_time_stream": (_I, [_P, _P, _P, _I, C.POINTER(_D)]),
    "heat_dot_static": (_I, [_P, _P, _I64, _P, _P]),
    "heat_set_inner_solver": (_I, [_P, _I]),
    "heat_set_pcg_geometry": (_I, [_P, _I, _I, _I]),
    "heat_pcg_geometry": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "heat_autotune_pcg": (_I, [_P, _I, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_D)]),
    "heat_batch_set_geometry": (_I, [_P, _I, _I]),
    "heat_batch_geometry": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "heat_batch_autotune": (_I, [_P, _I, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_D)]),
}
PCG_GEOMETRY = {"fused": 0, "one_launch": 1}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libheat_hip.so and declare every C-ABI signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"HIP library not built: {path} (run python -m mnhack8_amd.build)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise HeatError(rc, load_library().heat_last_error().decode())


class HeatOperator:
    """One mesh on the current HIP device (HeatOperator::setup)."""

    def __init__(self, n, dt, k, rhoc, kernel="fast", scatter="colored", verify=False,
                 no_source=False, source="gaussian", source_value=0.0, detect=False, _borrow=None):
        lib = load_library()
        self._lib = lib
        if _borrow is not None:                 # a handle owned by a HeatBatch
            self.h, self._owned = _borrow, False
            self.n, self.dt, self.detect = n, float(dt), bool(detect)
            nn, ne = _I64(), _I64()
            _check(lib.heat_sizes(self.h, C.byref(nn), C.byref(ne)))
            self.nn, self.ne = nn.value, ne.value
            return
        self._owned = True
        k = np.ascontiguousarray(k, dtype=np.float64).ravel()
        rhoc = np.ascontiguousarray(rhoc, dtype=np.float64).ravel()
        if k.size != n ** 3 or rhoc.size != n ** 3:
            raise ValueError(f"expected {n ** 3} element values")
        cfg = HeatConfig(n, float(dt), KERNEL[kernel], SCATTER[scatter], int(verify), int(no_source),
                         SOURCE[source], float(source_value), int(detect))
        h = C.c_void_p()
        _check(lib.heat_setup(C.byref(cfg), k.ctypes.data, rhoc.ctypes.data, C.byref(h)))
        self.h = h
        self.n, self.dt, self.detect = n, float(dt), bool(detect)
        nn, ne = _I64(), _I64()
        _check(lib.heat_sizes(h, C.byref(nn), C.byref(ne)))
        self.nn, self.ne = nn.value, ne.value

    # -- lifetime ----------------------------------------------------------
    def cleanup(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                _check(self._lib.heat_cleanup(self.h))
            self.h = None

    close = cleanup

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.cleanup()

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass

    # -- vectors -----------------------------------------------------------
    def ptr(self, v):
        if isinstance(v, str):
            p = C.c_void_p()
            _check(self._lib.heat_vector(self.h, VEC[v], C.byref(p)))
            return p.value
        if hasattr(v, "data_ptr"):
            import torch
            torch.cuda.synchronize()
            return v.data_ptr()
        return int(v)

    def upload(self, dst, host):
        a = np.ascontiguousarray(host, dtype=np.float64).ravel()
        _check(self._lib.heat_upload(self.h, self.ptr(dst), a.ctypes.data, a.size))

    def download(self, src, count=None):
        count = self.nn if count is None else count
        out = np.empty(count, dtype=np.float64)
        _check(self._lib.heat_download(self.h, out.ctypes.data, self.ptr(src), count))
        return out

    def synchronize(self):
        _check(self._lib.heat_synchronize(self.h))

    def field_stats(self):
        out = (C.c_double * 4)()
        _check(self._lib.heat_field_stats(self.h, out))
        return tuple(out)

    def set_colored_impl(self, impl):
        _check(self._lib.heat_set_colored_impl(self.h, COLORED_IMPL[impl]))

    def set_atomic_impl(self, impl):
        """--scatter atomic as "tiled" (default) or "element" (the reference's kernel shape)."""
        _check(self._lib.heat_set_atomic_impl(self.h, ATOMIC_IMPL[impl]))

    def set_march_chunk(self, planes):
        _check(self._lib.heat_set_march_chunk(self.h, int(planes)))

    def autotune(self, reps=20):
        """Time every colour-ordered schedule on this mesh and keep the fastest."""
        impl, planes, shape, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        _check(self._lib.heat_autotune(self.h, int(reps), C.byref(impl), C.byref(planes), C.byref(shape),
                                       C.byref(ms)))
        names = {v: k for k, v in COLORED_IMPL.items()}
        return names[impl.value], planes.value, TILE_SHAPES[shape.value], ms.value

    def set_tile_shape(self, shape):
        _check(self._lib.heat_set_tile_shape(self.h, int(shape)))

    def set_inner_solver(self, mode):
        """Inner solve of the refined solve: "one_launch" (default) or "standard"."""
        _check(self._lib.heat_set_inner_solver(self.h, {"one_launch": 0, "standard": 1}[mode]))

    def set_pcg_geometry(self, which, planes=0, wave=768):
        """Tile geometry of the PCG kernels ("fused" or "one_launch"): target
        planes per chunk (0: the default) and whole-wave fill (0: none)."""
        _check(self._lib.heat_set_pcg_geometry(self.h, PCG_GEOMETRY[which], int(planes), int(wave)))

    def pcg_geometry(self, which="fused"):
        """(planes, wave, chunks per column, tiles) of geometry `which`."""
        v = [C.c_int() for _ in range(4)]
        _check(self._lib.heat_pcg_geometry(self.h, PCG_GEOMETRY[which], *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune_pcg(self, which="fused", iters=30):
        """Time `iters` iterations per candidate geometry and keep the fastest:
        returns (planes, wave, ms per iteration)."""
        p, w, ms = C.c_int(), C.c_int(), C.c_double()
        _check(self._lib.heat_autotune_pcg(self.h, PCG_GEOMETRY[which], int(iters), C.byref(p), C.byref(w),
                                           C.byref(ms)))
        return p.value, w.value, ms.value

    def set_pcg_mode(self, mode):
        """"fused" (three launches per iteration) or "reference" (the reference's dot order)."""
        _check(self._lib.heat_set_pcg_mode(self.h, {"fused": 0, "reference": 1}[mode]))

    def apply_mrhs(self, X, Y, dt_eff):
        """Y[:, r] = P(rho c M + dt_eff kappa K) X[:, r] for 16 columns (FP32
        device arrays of shape (nn, 16), node-major)."""
        _check(self._lib.heat_apply_mrhs(self.h, self.ptr(X), self.ptr(Y), float(dt_eff)))

    def time_mrhs(self, X, Y, reps=20):
        ms = C.c_double()
        _check(self._lib.heat_time_mrhs(self.h, self.ptr(X), self.ptr(Y), int(reps), C.byref(ms)))
        return ms.value

    def arm_pcg_fault(self, vec, node, bit, iteration):
        """One-shot flip of bit `bit` of node `node` of PCG vector `vec`
        ("u", "r", "p", "z", "Ap") where iteration `iteration` produces it;
        for the refined solve also its FP64 state "x64" and "Ax64", where
        outer pass `iteration` stores it."""
        _check(self._lib.heat_arm_pcg_fault(self.h, SEU_VEC[vec], int(node), int(bit), int(iteration)))

    def pcg_fault_fired(self):
        f = C.c_int()
        _check(self._lib.heat_pcg_fault_fired(self.h, C.byref(f)))
        return bool(f.value)

    def mixed_status(self):
        """How the last refined solve ended: 1 tolerance met, 2 outer residual
        not finite, 0 pass limit."""
        f = C.c_int()
        _check(self._lib.heat_pcg_mixed_status(self.h, C.byref(f)))
        return f.value

    def set_tuning(self, pcg_chunk=8, use_graph=True):
        _check(self._lib.heat_set_tuning(self.h, int(pcg_chunk), int(bool(use_graph))))

    # -- HeatOperator methods ----------------------------------------------
    def apply_operator(self, x, y, dt_eff, allow_fi=False, project=True):
        _check(self._lib.heat_apply_operator(self.h, self.ptr(x), self.ptr(y), float(dt_eff),
                                             int(allow_fi), int(project)))

    def apply_operator_f32(self, x, y, dt_eff, project=True):
        _check(self._lib.heat_apply_operator_f32(self.h, self.ptr(x), self.ptr(y), float(dt_eff), int(project)))

    def apply_matvec(self, x, y):
        _check(self._lib.heat_apply_matvec(self.h, self.ptr(x), self.ptr(y)))

    def apply_preconditioner(self, r, z):
        _check(self._lib.heat_apply_preconditioner(self.h, self.ptr(r), self.ptr(z)))

    def build_rhs(self, Tn, b):
        _check(self._lib.heat_build_rhs(self.h, self.ptr(Tn), self.ptr(b)))

    def dot(self, x, y):
        out = C.c_double()
        _check(self._lib.heat_dot(self.h, self.ptr(x), self.ptr(y), C.byref(out)))
        return out.value

    def norm2(self, x):
        out = C.c_double()
        _check(self._lib.heat_norm2(self.h, self.ptr(x), C.byref(out)))
        return out.value

    def minmax(self, x):
        a, b = C.c_double(), C.c_double()
        _check(self._lib.heat_minmax(self.h, self.ptr(x), C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_dirichlet(self, v, value):
        _check(self._lib.heat_set_dirichlet(self.h, self.ptr(v), float(value)))

    def copy(self, x, y):
        _check(self._lib.heat_copy(self.h, self.ptr(x), self.ptr(y)))

    def axpy(self, a, x, y):
        _check(self._lib.heat_axpy(self.h, float(a), self.ptr(x), self.ptr(y)))

    def initial_condition(self, v, mode, t=0.0, alpha=1.0):
        modes = {"zero": 0, "gaussian": 1, "sine": 2}
        _check(self._lib.heat_initial_condition(self.h, self.ptr(v), modes.get(mode, mode), float(t), float(alpha)))

    # -- fault injection and detectors --------------------------------------
    def arm_fault(self, level, target, bit):
        _check(self._lib.heat_arm_fault(self.h, int(level), int(target), int(bit)))

    def disarm_fault(self):
        _check(self._lib.heat_disarm_fault(self.h))

    def flip_bit(self, v, idx, bit):
        _check(self._lib.heat_flip_bit(self.h, self.ptr(v), int(idx), int(bit)))

    def abft_threshold(self):
        """The D1 threshold chk_tol of the FP64 operator checksum."""
        t = C.c_double()
        _check(self._lib.heat_abft_threshold(self.h, C.byref(t)))
        return t.value

    def abft_take(self):
        m, c = C.c_double(), _I64()
        _check(self._lib.heat_abft_take(self.h, C.byref(m), C.byref(c)))
        return m.value, c.value

    def stream(self):
        p = C.c_void_p()
        _check(self._lib.heat_stream(self.h, C.byref(p)))
        return p.value

    def time_matvec(self, x, y, reps, flush_bytes=0):
        """Mean and minimum milliseconds per apply_matvec, timed on the device."""
        a, m = C.c_double(), C.c_double()
        _check(self._lib.heat_time_matvec(self.h, self.ptr(x), self.ptr(y), int(reps), int(flush_bytes),
                                          C.byref(a), C.byref(m)))
        return a.value, m.value

    def time_stream(self, x, y, reps):
        """Milliseconds per launch of a plain stream of the FP64 apply's
        algorithmic bytes (x, the two coefficient arrays read; y written)."""
        a = C.c_double()
        _check(self._lib.heat_time_stream(self.h, self.ptr(x), self.ptr(y), int(reps), C.byref(a)))
        return a.value

    # -- solvers --------------------------------------------------------------
    def pcg_solve(self, b, x, max_iter=500, tol=1e-8, history=False):
        it = C.c_int()
        hist = np.empty(3 * (max_iter + 1)) if history else None
        _check(self._lib.heat_pcg_solve(self.h, self.ptr(b), self.ptr(x), int(max_iter), float(tol),
                                        C.byref(it), hist.ctypes.data if history else None))
        if history:
            return it.value, hist.reshape(-1, 3)[: it.value + 1]
        return it.value

    def pcg_solve_mixed(self, b, x, max_outer=50, max_inner=500, tol=1e-8, tol_inner=1e-4):
        outer, inner = C.c_int(), C.c_int()
        _check(self._lib.heat_pcg_solve_mixed(self.h, self.ptr(b), self.ptr(x), int(max_outer), int(max_inner),
                                              float(tol), float(tol_inner), C.byref(outer), C.byref(inner)))
        return outer.value, inner.value


class HeatBatch:
    """An ensemble of independent solves advanced together (heat_batch_*).

    members: list of dicts with keys n, dt, k, rhoc and optionally source,
    source_value, no_source, verify, detect (all members or none).  ops[i] is
    a HeatOperator view of solve i (upload T^0, download T, arm faults, read
    its D1 record)."""

    def __init__(self, members):
        lib = load_library()
        self._lib = lib
        count = len(members)
        cfgs = (HeatConfig * count)()
        self._keep = []
        kp = (C.c_void_p * count)()
        rp = (C.c_void_p * count)()
        for i, m in enumerate(members):
            n = int(m["n"])
            k = np.ascontiguousarray(m["k"], dtype=np.float64).ravel()
            r = np.ascontiguousarray(m["rhoc"], dtype=np.float64).ravel()
            if k.size != n ** 3 or r.size != n ** 3:
                raise ValueError(f"member {i}: expected {n ** 3} element values")
            self._keep += [k, r]
            kp[i], rp[i] = k.ctypes.data, r.ctypes.data
            cfgs[i] = HeatConfig(n, float(m["dt"]), KERNEL["fast"], SCATTER["colored"], int(m.get("verify", 0)),
                                 int(m.get("no_source", 0)), SOURCE[m.get("source", "gaussian")],
                                 float(m.get("source_value", 0.0)), int(bool(m.get("detect", False))))
        h = C.c_void_p()
        _check(lib.heat_batch_setup(count, cfgs, kp, rp, C.byref(h)))
        self.h = h
        self.count = count
        self.ops = []
        for i, m in enumerate(members):
            oh = C.c_void_p()
            _check(lib.heat_batch_instance(h, i, C.byref(oh)))
            self.ops.append(HeatOperator(int(m["n"]), m["dt"], None, None, _borrow=oh))
        self._keep = []

    @classmethod
    def shared(cls, op, F):
        """COUNT right-hand sides on the mesh of HeatOperator `op` (kept by the
        caller).  F: (count, nn) source vectors, host array or device tensor."""
        self = cls.__new__(cls)
        self._lib = op._lib
        self.mesh = op
        F = np.ascontiguousarray(F, dtype=np.float64) if not hasattr(F, "data_ptr") else F
        count = F.shape[0]
        if hasattr(F, "data_ptr"):
            fptr = op.ptr(F)
        else:
            import torch
            Fd = torch.from_numpy(F).cuda()
            torch.cuda.synchronize()
            fptr = Fd.data_ptr()
        h = C.c_void_p()
        _check(self._lib.heat_batch_setup_rhs(op.h, count, fptr, C.byref(h)))
        self.h = h
        self.count = count
        self.ops = []
        return self

    def vector(self, i, which="T"):
        """Device pointer (int) of vector `which` of member i."""
        p = C.c_void_p()
        _check(self._lib.heat_batch_vector(self.h, int(i), VEC[which], C.byref(p)))
        return p.value

    def set_active(self, mask):
        a = np.ascontiguousarray(np.asarray(mask, dtype=np.int32))
        _check(self._lib.heat_batch_set_active(self.h, a.ctypes.data))

    def set_tuning(self, pcg_chunk=8, use_graph=True):
        _check(self._lib.heat_batch_set_tuning(self.h, int(pcg_chunk), int(bool(use_graph))))

    def set_geometry(self, planes=0, wave=768):
        """Fused tile geometry of every member (see HeatOperator.set_pcg_geometry)."""
        _check(self._lib.heat_batch_set_geometry(self.h, int(planes), int(wave)))

    def geometry(self, i):
        """(planes, wave, chunks per column, tiles) of member i."""
        v = [C.c_int() for _ in range(4)]
        _check(self._lib.heat_batch_geometry(self.h, int(i), *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, mode="fp64", iters=20):
        """Time one step of `iters` iterations per candidate geometry ("fp64"
        or "refined" steps), keep the fastest; every T is restored.  Returns
        (planes, wave, ms per iteration)."""
        p, w, ms = C.c_int(), C.c_int(), C.c_double()
        _check(self._lib.heat_batch_autotune(self.h, {"fp64": 0, "refined": 1}[mode], int(iters), C.byref(p),
                                             C.byref(w), C.byref(ms)))
        return p.value, w.value, ms.value

    def set_persistent(self, on=True):
        """Each batched solve as one persistent (cooperative) launch."""
        _check(self._lib.heat_batch_set_persistent(self.h, int(bool(on))))

    def detect(self):
        """Per-member detectors after a step of a batch with detectors:
        (d1, d2, tmin, tmax) arrays (NaN for masked members)."""
        out = [np.empty(self.count) for _ in range(4)]
        _check(self._lib.heat_batch_detect(self.h, *[o.ctypes.data for o in out]))
        return tuple(out)

    def step(self, nsteps=1, max_iter=500, tol=1e-8):
        """nsteps backward-Euler steps of every active solve; returns the
        (nsteps, count) iteration counts (-1: masked)."""
        it = np.zeros((max(nsteps, 1), self.count), dtype=np.int32)
        _check(self._lib.heat_batch_step(self.h, int(nsteps), int(max_iter), float(tol), it.ctypes.data))
        return it[:nsteps]

    def step_mixed(self, nsteps=1, max_outer=50, max_inner=500, tol=1e-8, tol_inner=1e-4):
        """nsteps backward-Euler steps of every active solve, each an
        FP32-refined solve (HeatOperator.pcg_solve_mixed); returns the
        (nsteps, count) inner-iteration totals (-1: masked)."""
        it = np.zeros((max(nsteps, 1), self.count), dtype=np.int32)
        _check(self._lib.heat_batch_step_mixed(self.h, int(nsteps), int(max_outer), int(max_inner), float(tol),
                                               float(tol_inner), it.ctypes.data))
        return it[:nsteps]

    def step_mrhs(self, nsteps=1, max_outer=50, max_inner=500, tol=1e-8, tol_inner=1e-4):
        """nsteps backward-Euler steps of a shared-mesh batch of 16 right-hand
        sides solved together: FP32 inner PCG with the matrix-core operator in
        FP64 defect correction.  Returns the (nsteps, 16) inner-iteration
        totals and the (nsteps, 16) largest D1 ratios (0 without detectors)."""
        it = np.zeros((max(nsteps, 1), self.count), dtype=np.int32)
        d1 = np.zeros((max(nsteps, 1), self.count), dtype=np.float64)
        _check(self._lib.heat_batch_step_mrhs(self.h, int(nsteps), int(max_outer), int(max_inner), float(tol),
                                              float(tol_inner), it.ctypes.data, d1.ctypes.data))
        return it[:nsteps], d1[:nsteps]

    def cleanup(self):
        if getattr(self, "h", None):
            for op in self.ops:
                op.h = None
            _check(self._lib.heat_batch_cleanup(self.h))
            self.h = None

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass


def pcg_solve(op: HeatOperator, b, x, max_iter=500, tol=1e-8, history=False):
    """pcg_solve(op, b, x, max_iter, tol) of src/heat_solver_het.cu:994."""
    return op.pcg_solve(b, x, max_iter, tol, history)
