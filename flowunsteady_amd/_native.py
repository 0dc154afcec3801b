"""ctypes binding of libvpmx.so (the C ABI declared in include/vpmx.h).

The argument lists here are exactly what a Julia `ccall` passes: plain
pointers, 64-bit counts and the two small parameter structs.  There is no
fallback: if the built library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvpmx.so")

# status codes
OK, EINVAL, ENOMEM, EHIP, ENODEV, EINJECT, EUNSUPPORTED = 0, 1, 2, 3, 4, 5, 6
# kernels
SINGULAR, GAUSSIAN, GAUSSIANERF, WINCKELMANS = 0, 1, 2, 3
# precision
FP64, MIXED = 0, 1
# relaxation
RELAX_NONE, RELAX_PEDRIZZETTI, RELAX_CORRECTED = 0, 1, 2
# SFS
SFS_NONE, SFS_CONSTANT, SFS_DYNAMIC = 0, 1, 2
# timing classes
T_UJ, T_ESTR, T_UPDATE, T_FMM, T_GRAPH, T_DSIG = 0, 1, 2, 3, 4, 5
# schedules of the batched pairwise pass
SCHED_FLAT, SCHED_QUEUE = 0, 1
# viscous schemes
VISCOUS_NONE, VISCOUS_CORESPREADING, VISCOUS_PSE = 0, 1, 2
FMM_STATS = ("ms_tree", "ms_traverse", "ms_upload", "cells", "leaves", "m2l_pairs",
             "p2p_leaf_pairs", "p2p_pairs", "ms_eval")


class Layout(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int32), ("X", ctypes.c_int32), ("Gamma", ctypes.c_int32),
                ("sigma", ctypes.c_int32), ("vol", ctypes.c_int32), ("U", ctypes.c_int32),
                ("J", ctypes.c_int32), ("SFS", ctypes.c_int32), ("C", ctypes.c_int32),
                ("static_", ctypes.c_int32)]


class UJParams(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("transposed", ctypes.c_int32), ("reset", ctypes.c_int32),
                ("sfs", ctypes.c_int32), ("fmm", ctypes.c_int32), ("p", ctypes.c_int32),
                ("ncrit", ctypes.c_int32), ("theta", ctypes.c_double),
                ("sigma_cut", ctypes.c_double)]


class StepParams(ctypes.Structure):
    _fields_ = [("f", ctypes.c_double), ("g", ctypes.c_double), ("Uinf", ctypes.c_double * 3),
                ("sfs", ctypes.c_int32), ("Cs", ctypes.c_double),
                ("clip_backscatter", ctypes.c_int32), ("relax_kind", ctypes.c_int32),
                ("rlxf", ctypes.c_double), ("alpha", ctypes.c_double),
                ("sfs_rlxf", ctypes.c_double), ("minC", ctypes.c_double),
                ("maxC", ctypes.c_double), ("viscous", ctypes.c_int32), ("nu", ctypes.c_double),
                ("sgm0", ctypes.c_double), ("beta", ctypes.c_double), ("itmax", ctypes.c_int32),
                ("tol", ctypes.c_double)]


class VPMXError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vpmx error {code}: {msg}")
        self.code = code


_dp = ctypes.POINTER(ctypes.c_double)
_ctxp = ctypes.c_void_p
_dpp = ctypes.POINTER(_dp)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes); every function the header declares
SIGNATURES = {
    "vpmx_create": (ctypes.c_int, [ctypes.POINTER(_ctxp), ctypes.c_int32]),
    "vpmx_destroy": (None, [_ctxp]),
    "vpmx_last_error": (ctypes.c_char_p, [_ctxp]),
    "vpmx_version": (ctypes.c_char_p, []),
    "vpmx_uj": (ctypes.c_int, [_ctxp, _dp, ctypes.c_int64, ctypes.POINTER(Layout),
                               ctypes.POINTER(UJParams)]),
    "vpmx_uj_device_noalloc": (ctypes.c_int, [_ctxp, ctypes.c_int64, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.POINTER(UJParams),
                                              ctypes.c_void_p]),
    "vpmx_rk3_step": (ctypes.c_int, [_ctxp, _dp, ctypes.c_int64, ctypes.POINTER(Layout),
                                     ctypes.POINTER(UJParams), ctypes.POINTER(StepParams),
                                     ctypes.c_double, ctypes.c_int32]),
    "vpmx_euler_step": (ctypes.c_int, [_ctxp, _dp, ctypes.c_int64, ctypes.POINTER(Layout),
                                       ctypes.POINTER(UJParams), ctypes.POINTER(StepParams),
                                       ctypes.c_double, ctypes.c_int32]),
    "vpmx_upload": (ctypes.c_int, [_ctxp, _dp, ctypes.c_int64, ctypes.POINTER(Layout)]),
    "vpmx_download": (ctypes.c_int, [_ctxp, _dp, ctypes.c_int64, ctypes.POINTER(Layout)]),
    "vpmx_uj_resident": (ctypes.c_int, [_ctxp, ctypes.POINTER(UJParams)]),
    "vpmx_rk3_resident": (ctypes.c_int, [_ctxp, ctypes.POINTER(UJParams),
                                         ctypes.POINTER(StepParams), ctypes.c_double,
                                         ctypes.c_int32]),
    "vpmx_euler_resident": (ctypes.c_int, [_ctxp, ctypes.POINTER(UJParams),
                                           ctypes.POINTER(StepParams), ctypes.c_double,
                                           ctypes.c_int32]),
    "vpmx_diagnostics": (ctypes.c_int, [_ctxp, _dp, _dp]),
    "vpmx_uj_targets": (ctypes.c_int, [_ctxp, _dp, ctypes.c_int64, _dp, _dp,
                                       ctypes.POINTER(UJParams)]),
    "vpmx_resident_np": (ctypes.c_int64, [_ctxp]),
    "vpmx_set_timing": (ctypes.c_int, [_ctxp, ctypes.c_int32]),
    "vpmx_get_timing": (ctypes.c_int, [_ctxp, ctypes.c_int32, _dp,
                                       ctypes.POINTER(ctypes.c_int64)]),
    "vpmx_stream": (ctypes.c_void_p, [_ctxp]),
    "vpmx_fmm_stats": (ctypes.c_int, [_ctxp, _dp, ctypes.c_int32]),
    "vpmx_fmm_plan_counts": (ctypes.c_int, [ctypes.c_int64, _dp, _dp, ctypes.c_int32,
                                            ctypes.c_double, ctypes.c_double,
                                            ctypes.POINTER(ctypes.c_int64)]),
    "vpmx_inject_fault": (ctypes.c_int, [_ctxp, ctypes.c_int32]),
    "vpmx_batch_upload": (ctypes.c_int, [_ctxp, ctypes.c_int32, _dpp, _i64p,
                                         ctypes.POINTER(Layout)]),
    "vpmx_batch_download": (ctypes.c_int, [_ctxp, ctypes.c_int32, _dpp, _i64p,
                                           ctypes.POINTER(Layout)]),
    "vpmx_batch_uj": (ctypes.c_int, [_ctxp, ctypes.c_int32, _dpp, _i64p, ctypes.POINTER(Layout),
                                     ctypes.POINTER(UJParams)]),
    "vpmx_batch_rk3_step": (ctypes.c_int, [_ctxp, ctypes.c_int32, _dpp, _i64p,
                                           ctypes.POINTER(Layout), ctypes.POINTER(UJParams),
                                           ctypes.POINTER(StepParams), _dp, ctypes.c_double,
                                           ctypes.c_int32]),
    "vpmx_batch_rk3_resident": (ctypes.c_int, [_ctxp, ctypes.POINTER(UJParams),
                                               ctypes.POINTER(StepParams), _dp, ctypes.c_double,
                                               ctypes.c_int32]),
    "vpmx_batch_nfields": (ctypes.c_int32, [_ctxp]),
    "vpmx_batch_set_schedule": (ctypes.c_int, [_ctxp, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32]),
    "vpmx_batch_get_schedule": (ctypes.c_int, [_ctxp, _i32p, _i32p, _i32p]),
    "vpmx_batch_autotune": (ctypes.c_int, [_ctxp, ctypes.POINTER(UJParams), ctypes.c_int32, _dp,
                                           ctypes.c_int32, _i32p]),
    "vpmx_set_graph": (ctypes.c_int, [_ctxp, ctypes.c_int32]),
    "vpmx_core_spreading_reset": (ctypes.c_int, [_ctxp, ctypes.POINTER(UJParams),
                                                 ctypes.POINTER(StepParams), ctypes.c_int32,
                                                 _i32p, _i32p, _dp]),
    "vpmx_ffv": (ctypes.c_int, [_ctxp, _dp, ctypes.c_int64, ctypes.POINTER(Layout), ctypes.c_int32,
                                _dp]),
    "vpmx_radial_f32": (ctypes.c_int, [ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                       ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libvpmx.so; raise if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "`python -m flowunsteady_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error(ctx=None) -> str:
    m = lib().vpmx_last_error(ctx)
    return m.decode() if m else ""


def check(code: int, ctx=None):
    if code != OK:
        raise VPMXError(code, last_error(ctx))


class Context:
    """Owns one vpmx_ctx (device buffers, stream, timers)."""

    def __init__(self, device: int = 0):
        p = _ctxp()
        check(lib().vpmx_create(ctypes.byref(p), int(device)))
        self.ptr = p
        self.device = device

    def close(self):
        if getattr(self, "ptr", None):
            lib().vpmx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, code):
        check(code, self.ptr)

    def set_timing(self, on: bool):
        self.check(lib().vpmx_set_timing(self.ptr, int(on)))

    def timing(self, which: int):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self.check(lib().vpmx_get_timing(self.ptr, which, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def stream(self) -> int:
        return lib().vpmx_stream(self.ptr)

    def set_graph(self, on: bool):
        self.check(lib().vpmx_set_graph(self.ptr, int(on)))

    def set_schedule(self, schedule: int, block: int, grid: int = 0):
        self.check(lib().vpmx_batch_set_schedule(self.ptr, schedule, block, grid))

    def schedule(self) -> tuple:
        a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self.check(lib().vpmx_batch_get_schedule(self.ptr, ctypes.byref(a), ctypes.byref(b),
                                                 ctypes.byref(c)))
        return a.value, b.value, c.value

    def autotune(self, params, reps: int = 1) -> list:
        """Times every candidate schedule of the resident batch; keeps the fastest.
        Returns [(schedule, block, ms), ...] in candidate order."""
        ms = (ctypes.c_double * 8)()
        n = ctypes.c_int32()
        self.check(lib().vpmx_batch_autotune(self.ptr, ctypes.byref(params), int(reps), ms, 8,
                                             ctypes.byref(n)))
        if params.precision == MIXED:
            cands = [(s, b) for b in (64, 128, 256) for s in (SCHED_FLAT, SCHED_QUEUE)]
        else:
            cands = [(SCHED_FLAT, b) for b in (64, 128, 256)]
        return [(s, b, ms[i]) for i, (s, b) in enumerate(cands[:n.value])]

    def fmm_stats(self) -> dict:
        out = (ctypes.c_double * len(FMM_STATS))()
        self.check(lib().vpmx_fmm_stats(self.ptr, out, len(FMM_STATS)))
        return dict(zip(FMM_STATS, list(out)))


def pointer_array(mats) -> "ctypes.Array":
    """double*[] of the column-major float64 matrices in `mats` (kept alive by the caller)."""
    arr = (_dp * len(mats))()
    for k, m in enumerate(mats):
        arr[k] = m.ctypes.data_as(_dp)
    return arr


def fmm_plan_counts(X, sigma, ncrit=50, theta=0.4, sigma_cut=7.0) -> dict:
    """Tree and traversal counts of the library's FMM plan, built on the host."""
    import numpy as np
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
    s = np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1)
    out = (ctypes.c_int64 * 5)()
    check(lib().vpmx_fmm_plan_counts(X.shape[0], X.ctypes.data_as(_dp), s.ctypes.data_as(_dp),
                                     int(ncrit), float(theta), float(sigma_cut), out))
    return dict(zip(("ncell", "nleaf", "nm2l", "np2p", "p2p_pairs"), list(out)))
