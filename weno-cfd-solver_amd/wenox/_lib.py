"""ctypes binding of libwenox.so (C ABI declared in include/wenox.h)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwenox.so")

OK, E_ARG, E_SHAPE, E_DEVICE = 0, -1, -2, 1
EXACT, FAST, TWIN = 0, 1, 2


class WenoxError(RuntimeError):
    """A HIP runtime failure or a missing/unsupported device."""


class PanicException(BaseException):
    """Raised where the reference Rust extension panics (too-small arrays,
    fewer than 4 channels); like pyo3_runtime.PanicException it derives from
    BaseException."""


class View3(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p),
                ("dim", ctypes.c_int64 * 3),
                ("stride", ctypes.c_int64 * 3)]


class View2(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p),
                ("dim", ctypes.c_int64 * 2),
                ("stride", ctypes.c_int64 * 2)]


class SolverDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
                ("dx", ctypes.c_double), ("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
                ("cons_switch", ctypes.c_int), ("force_hlle", ctypes.c_int),
                ("jp_lo", ctypes.c_double), ("jp_hi", ctypes.c_double),
                ("bc", ctypes.c_int), ("dmr_bound", ctypes.c_int), ("dmr_x0", ctypes.c_double),
                ("engine", ctypes.c_int), ("variant", ctypes.c_int), ("precision", ctypes.c_int),
                ("use_graph", ctypes.c_int)]


_lib = None

DROPIN_SYMBOLS = [
    "wenox_flux_x", "wenox_flux_y", "wenox_con2primi", "wenox_weno_x", "wenox_weno_y",
    "wenox_hllc_x", "wenox_hllc_y", "wenox_hlle_x", "wenox_hlle_y", "wenox_hllc_hlle_x",
    "wenox_hllc_hlle_y", "wenox_l_local",
]


def load():
    """Load libwenox.so (raises WenoxError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise WenoxError(f"{LIB_PATH} is missing: build it with "
                         "`python weno-cfd-solver_amd/build.py` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    V = ctypes.POINTER(View3)
    d, i, dp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    sig = {
        "wenox_last_error": ([], ctypes.c_char_p),
        "wenox_version": ([], ctypes.c_char_p),
        "wenox_set_variant": ([i], i),
        "wenox_device_count": ([], i),
        "wenox_flux_x": ([V, d, dp], i),
        "wenox_flux_y": ([V, d, dp], i),
        "wenox_con2primi": ([V, d, dp, dp, dp, dp, dp, dp], i),
        "wenox_weno_x": ([V, d, i, dp, dp, dp, dp], i),
        "wenox_weno_y": ([V, d, i, dp, dp, dp, dp], i),
        "wenox_hllc_x": ([V, V, V, V, d, dp], i),
        "wenox_hllc_y": ([V, V, V, V, d, dp], i),
        "wenox_hlle_x": ([V, V, V, V, d, dp], i),
        "wenox_hlle_y": ([V, V, V, V, d, dp], i),
        "wenox_hllc_hlle_x": ([V, V, V, V, d, d, d, i, dp], i),
        "wenox_hllc_hlle_y": ([V, V, V, V, d, d, d, i, dp], i),
        "wenox_l_local": ([V, d, d, i, d, d, i, dp], i),
        "wenox_l_local_noalloc": ([V, d, d, i, d, d, i, dp, dp], i),
        "wenox_solver_create": ([ctypes.POINTER(SolverDesc), ctypes.POINTER(ctypes.c_void_p)], i),
        "wenox_solver_destroy": ([dp], None),
        "wenox_solver_upload": ([dp, V, d], i),
        "wenox_solver_download": ([dp, dp], i),
        "wenox_solver_run": ([dp, ctypes.c_int64, d, ctypes.POINTER(d),
                              ctypes.POINTER(ctypes.c_int64)], i),
        "wenox_solver_rhs": ([dp, dp], i),
        "wenox_solver_conserved_sums": ([dp, dp, i], i),
        "wenox_solver_last_ms_per_step": ([dp], d),
        "wenox_solver_profile": ([dp, i, dp, i], i),
        "wenox_solver_plan": ([dp, ctypes.POINTER(i), ctypes.POINTER(i)], i),
        "wenox_solver_set_chunks": ([dp, i, i], i),
        "wenox_solver_tiling": ([dp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)], i),
        "wenox_solver_set_tiling": ([dp, i, i, i, i], i),
        "wenox_solver_combine_cols": ([dp, ctypes.POINTER(i)], i),
        "wenox_solver_set_combine_cols": ([dp, i], i),
        "wenox_solver_pipeline": ([dp, ctypes.POINTER(i)], i),
        "wenox_solver_set_pipeline": ([dp, i], i),
        "wenox_solver_split_jump": ([dp, ctypes.POINTER(d)], i),
        "wenox_solver_set_split_jump": ([dp, d], i),
        "wenox_solver_layout": ([dp, ctypes.POINTER(i)], i),
        "wenox_solver_set_layout": ([dp, i], i),
        "wenox_solver_set_cfl_rule": ([dp, d, d], i),
        "wenox_solver_set_pfloor": ([dp, d], i),
        "wenox_objective_2d": ([ctypes.POINTER(View2)] * 3 + [d, d, dp, dp, dp], i),
        "wenox_solver_objective": ([dp, d, d, dp, dp, dp], i),
        "wenox_objective_fields": ([ctypes.POINTER(View2), d, d, dp, dp, dp, dp], i),
        "wenox_batch_create": ([ctypes.POINTER(SolverDesc), i, dp, dp, dp, i,
                                ctypes.POINTER(ctypes.c_void_p)], i),
        "wenox_batch_destroy": ([dp], None),
        "wenox_batch_upload": ([dp, i, V, d], i),
        "wenox_batch_download": ([dp, i, dp], i),
        "wenox_batch_run": ([dp, ctypes.c_int64, d, ctypes.POINTER(ctypes.c_int64)], i),
        "wenox_batch_chunks": ([dp, ctypes.POINTER(i), ctypes.POINTER(i)], i),
        "wenox_batch_set_chunks": ([dp, i, i], i),
        "wenox_batch_profile": ([dp, i, dp], i),
        "wenox_batch_layout": ([dp, ctypes.POINTER(i)], i),
        "wenox_batch_set_layout": ([dp, i], i),
        "wenox_batch_persistent": ([dp, ctypes.POINTER(i)], i),
        "wenox_batch_set_persistent": ([dp, i], i),
        "wenox_batch_set_spin_limit": ([dp, i], i),
        "wenox_batch_set_split_jump": ([dp, d], i),
        "wenox_batch_packs": ([dp, ctypes.POINTER(i), ctypes.POINTER(i)], i),
        "wenox_batch_set_pack_max": ([dp, i], i),
        "wenox_batch_persist_xcds": ([dp, ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint)], i),
        "wenox_batch_set_persist_xcds": ([dp, ctypes.c_uint], i),
        "wenox_batch_profile_steps": ([dp, i, dp], i),
        "wenox_batch_status": ([dp, dp, dp], i),
        "wenox_batch_last_ms_per_step": ([dp], d),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def last_error() -> str:
    return load().wenox_last_error().decode(errors="replace")


def check(rc: int) -> None:
    if rc == OK:
        return
    msg = last_error()
    if rc == E_SHAPE:
        raise PanicException(msg)
    if rc == E_ARG:
        raise ValueError(msg)
    raise WenoxError(msg)


def view2(a: np.ndarray) -> View2:
    """Describe a 2-D float64 array (non-multiple-of-8 strides get a
    contiguous copy, kept alive through the View2's _keep)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 2:
        raise TypeError("expected a 2-D array")
    if any(s % 8 for s in a.strides):
        a = np.ascontiguousarray(a)
    v = View2(a.ctypes.data, (ctypes.c_int64 * 2)(*a.shape), (ctypes.c_int64 * 2)(*[s // 8 for s in a.strides]))
    v._keep = a
    return v


def view(a: np.ndarray) -> View3:
    """Describe a float64 array; non-multiple-of-8 strides get a contiguous copy
    (the caller keeps the returned array alive through the View3's _keep)."""
    if any(s % 8 for s in a.strides):
        a = np.ascontiguousarray(a)
    v = View3(a.ctypes.data, (ctypes.c_int64 * 3)(*a.shape), (ctypes.c_int64 * 3)(*[s // 8 for s in a.strides]))
    v._keep = a
    return v
