"""ctypes binding of libparallelkde_hip (the C ABI in include/parallelkde_hip.h).

The library is required: importing the binding of a missing or unloadable
library raises immediately; nothing falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get(
    "PKDE_LIB", os.path.join(os.path.dirname(_HERE), "lib", "libparallelkde_hip.so"))
MAX_DIMS = 6
ABI_VERSION = 2

OK = 0
INVALID_ARGUMENT = 1
OUT_OF_GRID = 2
NO_DEVICE = 3
OUT_OF_MEMORY = 4
DEVICE_ERROR = 5
UNSUPPORTED = 6
PENDING = 7

EXPORTED_SYMBOLS = [
    "pkde_version", "pkde_last_error", "pkde_gradepro_params_default",
    "pkde_context_create", "pkde_context_destroy", "pkde_device_free_memory", "pkde_find_grid",
    "pkde_gradepro_plan_create", "pkde_gradepro_plan_destroy", "pkde_gradepro_plan_footprint",
    "pkde_rot_plan_create",
    "pkde_gradepro_estimate", "pkde_gradepro_estimate_device", "pkde_gradepro_check",
    "pkde_gradepro_last_stats",
    "pkde_gradepro_debug_enable", "pkde_gradepro_debug_vmr", "pkde_gradepro_convergence",
    "pkde_gradepro_graph_enable", "pkde_gradepro_autotune",
    "pkde_dirac_sequences", "pkde_fft_forward", "pkde_gradepro_estimate_batch",
    "pkde_context_release_batch",
    "pkde_gradepro_estimate_batch_device", "pkde_gradepro_estimate_1d_noalloc",
    "pkde_gradepro_request_footprint",
    "pkde_server_create", "pkde_server_submit", "pkde_server_poll", "pkde_server_wait",
    "pkde_server_stats_get", "pkde_server_destroy",
]


class Params(C.Structure):
    _fields_ = [
        ("n_bootstraps", C.c_int32),
        ("n_steps", C.c_int32),
        ("bw_step", C.c_double),
        ("bw_final", C.c_double * MAX_DIMS),
        ("alpha_s", C.c_double),
        ("alpha_os", C.c_double),
        ("eps", C.c_double),
        ("rng_seed", C.c_uint64),
        ("deterministic_binning", C.c_int32),
        ("precision", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("n_rungs", C.c_int32),
        ("steps_low", C.c_int32),
        ("steps_over", C.c_int32),
        ("eps", C.c_double),
        ("bw_final", C.c_double * MAX_DIMS),
        ("ms_total", C.c_double),
        ("ms_ladder", C.c_double),
        ("algorithmic_bytes", C.c_double),
    ]


class Request(C.Structure):
    _fields_ = [
        ("ndims", C.c_int32),
        ("n_per_axis", C.c_int64 * MAX_DIMS),
        ("lo", C.c_double * MAX_DIMS),
        ("hi", C.c_double * MAX_DIMS),
        ("h", C.c_double * MAX_DIMS),
        ("samples", C.c_void_p),
        ("n", C.c_int64),
        ("params", Params),
        ("density_out", C.c_void_p),
        ("chosen_out", C.c_void_p),
    ]


class ServerCfg(C.Structure):
    _fields_ = [("max_inflight", C.c_int32), ("memory_budget", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [
        ("completed", C.c_uint64), ("rejected", C.c_uint64), ("queued", C.c_uint64),
        ("inflight", C.c_uint64), ("bytes_in_use", C.c_uint64),
        ("p50_ms", C.c_double), ("p99_ms", C.c_double),
    ]


_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises OSError if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"libparallelkde_hip.so not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i64p, dp, i32p, i16p = C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double), \
        C.POINTER(C.c_int32), C.POINTER(C.c_int16)
    L.pkde_version.restype = C.c_int
    L.pkde_last_error.restype = C.c_char_p
    L.pkde_gradepro_params_default.argtypes = [C.POINTER(Params)]
    L.pkde_context_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.pkde_context_destroy.argtypes = [vp]
    L.pkde_device_free_memory.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.pkde_find_grid.argtypes = [vp, C.c_int, vp, C.c_int64, vp, vp, vp]
    L.pkde_gradepro_plan_create.argtypes = [vp, C.c_int, i64p, dp, dp, dp, C.POINTER(Params), C.POINTER(vp)]
    L.pkde_gradepro_plan_destroy.argtypes = [vp]
    L.pkde_rot_plan_create.argtypes = [vp, C.c_int, i64p, dp, dp, dp, C.c_int, C.c_int, C.POINTER(vp)]
    L.pkde_gradepro_plan_footprint.argtypes = [vp, C.c_int64]
    L.pkde_gradepro_plan_footprint.restype = C.c_uint64
    L.pkde_gradepro_estimate.argtypes = [vp, vp, C.c_int64, vp, vp, vp, C.POINTER(Stats)]
    L.pkde_gradepro_estimate_device.argtypes = [vp, vp, C.c_int64, vp, vp, vp, vp]
    L.pkde_gradepro_check.argtypes = [vp]
    L.pkde_gradepro_last_stats.argtypes = [vp, C.POINTER(Stats)]
    L.pkde_gradepro_debug_enable.argtypes = [vp, C.c_int]
    L.pkde_gradepro_graph_enable.argtypes = [vp, C.c_int]
    L.pkde_gradepro_autotune.argtypes = [vp, C.c_int32, vp, vp]
    L.pkde_gradepro_debug_vmr.argtypes = [vp, vp, C.c_int64]
    L.pkde_gradepro_convergence.argtypes = [vp, C.c_int64, C.c_int32, vp, vp, C.c_double, C.c_int32,
                                            C.c_int32, vp]
    L.pkde_dirac_sequences.argtypes = [vp, C.c_int, i64p, dp, dp, dp, vp, C.c_int64, vp, C.c_int32,
                                       C.c_int, vp, vp]
    L.pkde_fft_forward.argtypes = [vp, C.c_int32, vp, vp]
    L.pkde_gradepro_estimate_batch.argtypes = [vp, C.c_int32, C.POINTER(Request), vp]
    L.pkde_context_release_batch.argtypes = [vp]
    L.pkde_gradepro_estimate_1d_noalloc.argtypes = [vp, C.c_int64, C.c_double, C.c_double,
                                                    C.c_double, C.POINTER(Params), vp, C.c_int64,
                                                    vp, vp, vp]
    L.pkde_gradepro_estimate_batch_device.argtypes = [vp, C.c_int32, C.POINTER(Request), vp]
    L.pkde_gradepro_request_footprint.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.pkde_server_create.argtypes = [vp, C.POINTER(ServerCfg), C.POINTER(vp)]
    L.pkde_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.pkde_server_poll.argtypes = [vp, C.c_uint64, C.POINTER(C.c_int)]
    L.pkde_server_wait.argtypes = [vp, C.c_uint64]
    L.pkde_server_stats_get.argtypes = [vp, C.POINTER(ServerStats)]
    L.pkde_server_destroy.argtypes = [vp]
    if L.pkde_version() != ABI_VERSION:
        raise OSError("libparallelkde_hip ABI version mismatch")
    _lib = L
    return L


def last_error() -> str:
    msg = load().pkde_last_error()
    return msg.decode() if msg else ""


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data
