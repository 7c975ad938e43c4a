"""ctypes binding of the FP64 CPU oracle (oracle/libgradepro_oracle.so).

Test infrastructure: the oracle is the checker the HIP path is compared
against.  Only tests, the smoke check and bench.py's CPU-baseline leg import
this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "libgradepro_oracle.so")
MAX_DIMS = 6

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class OgGrid(C.Structure):
    _fields_ = [
        ("ndims", C.c_int),
        ("n", C.c_int64 * MAX_DIMS),
        ("lo", C.c_double * MAX_DIMS),
        ("hi", C.c_double * MAX_DIMS),
        ("h", C.c_double * MAX_DIMS),
    ]


class OgParams(C.Structure):
    _fields_ = [
        ("n_bootstraps", C.c_int32),
        ("n_steps", C.c_int32),
        ("bw_step", C.c_double),
        ("bw_final", C.c_double * MAX_DIMS),
        ("alpha_s", C.c_double),
        ("alpha_os", C.c_double),
        ("eps", C.c_double),
        ("threads", C.c_int32),
    ]


class OgInfo(C.Structure):
    _fields_ = [
        ("n_rungs", C.c_int32),
        ("steps_low", C.c_int32),
        ("steps_over", C.c_int32),
        ("eps", C.c_double),
        ("bw_final", C.c_double * MAX_DIMS),
        ("seconds_binning_fft", C.c_double),
        ("seconds_ladder", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.og_params_default.argtypes = [C.POINTER(OgParams)]
        L.og_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.og_bootstrap_indices.argtypes = [C.c_uint64, C.c_int64, C.c_int32, _i32p]
        L.og_fftfreq.argtypes = [C.c_int64, C.c_double, _dp]
        L.og_silverman.argtypes = [C.c_int, C.c_int64, _dp, _dp]
        L.og_ladder.argtypes = [C.c_int, _dp, C.c_int32, C.c_double, C.c_void_p, C.c_void_p]
        L.og_ladder.restype = C.c_int
        L.og_duration_steps.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double,
                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.og_dlogt.argtypes = [C.c_int, _dp, _dp, C.c_void_p]
        L.og_dlogt.restype = C.c_double
        L.og_check_data.argtypes = [C.POINTER(OgGrid), _dp, C.c_int64]
        L.og_dirac_sequences.argtypes = [C.POINTER(OgGrid), _dp, C.c_int64, _i32p, C.c_int32,
                                         _dp, C.c_void_p, C.c_int]
        L.og_fft_nd.argtypes = [C.c_int, C.POINTER(C.c_int64), _dp, C.c_int]
        L.og_set_fft_order.argtypes = [C.c_int]
        L.og_set_deposit_order.argtypes = [C.c_int]
        L.og_set_real_output.argtypes = [C.c_int]
        L.og_fixed_frac_bits.argtypes = [C.c_int64]
        L.og_fixed_frac_bits.restype = C.c_int
        L.og_propagate.argtypes = [C.POINTER(OgGrid), _dp, _dp, _dp, _dp, C.c_void_p, _dp, C.c_void_p]
        L.og_scaled_vmr.argtypes = [C.c_int, C.c_int64, C.c_int32, _dp, _dp, _dp, _dp, C.c_int64, _dp]
        L.og_full_means.argtypes = [C.c_int64, _dp, _dp]
        L.og_identify_convergence.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                              C.c_double, C.c_int, C.c_int, _dp, _u16p, _u16p,
                                              _u8p, C.c_void_p, C.c_int16, C.c_int]
        L.og_gradepro_estimate.argtypes = [C.POINTER(OgGrid), _dp, C.c_int64, C.c_void_p, C.c_uint64,
                                           C.POINTER(OgParams), _dp, _i16p, C.c_void_p, C.c_void_p,
                                           C.c_int32, C.POINTER(OgInfo)]
        L.og_gradepro_estimate.restype = C.c_int
        L.og_rot_estimate.argtypes = [C.POINTER(OgGrid), _dp, C.c_int64, C.c_int, _dp, _dp]
        L.og_rot_estimate.restype = C.c_int
        _lib = L
    return _lib


def make_grid(lo, hi, n) -> OgGrid:
    """Grid from per-axis (lo, hi, n); spacing (hi-lo)/(n-1) like Julia's
    range(lo, hi, length=n) (Grids.jl:67-87)."""
    lo = np.atleast_1d(np.asarray(lo, dtype=np.float64))
    hi = np.atleast_1d(np.asarray(hi, dtype=np.float64))
    n = np.atleast_1d(np.asarray(n, dtype=np.int64))
    g = OgGrid()
    g.ndims = len(n)
    for i in range(len(n)):
        g.n[i] = int(n[i])
        g.lo[i] = float(lo[i])
        g.hi[i] = float(hi[i])
        g.h[i] = float((hi[i] - lo[i]) / (n[i] - 1))
    return g


def make_grid_step(lo, hi, h, n) -> OgGrid:
    """Grid with an explicit step (Julia colon ranges such as -5.0:0.1:5.0)."""
    g = OgGrid()
    g.ndims = len(n)
    for i in range(len(n)):
        g.n[i], g.lo[i], g.hi[i], g.h[i] = int(n[i]), float(lo[i]), float(hi[i]), float(h[i])
    return g


def grid_shape(g: OgGrid):
    return tuple(int(g.n[i]) for i in range(g.ndims))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().og_philox4x32_10(c, k, o)
    return [int(o[i]) for i in range(4)]


def bootstrap_indices(seed, n, B):
    idx = np.zeros(n * B, dtype=np.int32)
    lib().og_bootstrap_indices(seed, n, B, idx)
    return idx.reshape(B, n)  # [b][s]


def fftfreq(n, h):
    out = np.zeros(n)
    lib().og_fftfreq(n, h, out)
    return out


def silverman(samples_col):
    """samples_col: array (n, ndims) i.e. column-major ndims x n."""
    s = np.ascontiguousarray(samples_col, dtype=np.float64)
    n, d = s.shape
    out = np.zeros(d)
    lib().og_silverman(d, n, s.ravel(), out)
    return out


def fft_nd(arr: np.ndarray, sign: int) -> np.ndarray:
    """Unnormalised C2C over all axes of a column-major complex array given
    in numpy (C-order) with reversed axes: arr shape (n_N, ..., n_1)."""
    a = np.ascontiguousarray(arr, dtype=np.complex128)
    shape_cm = list(reversed(a.shape))
    buf = a.view(np.float64).ravel().copy()
    n = (C.c_int64 * len(shape_cm))(*shape_cm)
    lib().og_fft_nd(len(shape_cm), n, buf, sign)
    return buf.view(np.complex128).reshape(a.shape)


def set_fft_order(mode: int):
    """Radix order of the oracle's FFT (og_set_fft_order): 0 default,
    1 reversed, 2 radix-2 passes for the power of two."""
    lib().og_set_fft_order(int(mode))


def set_deposit_order(mode: int):
    """Deposit mode of the oracle's binning (og_set_deposit_order): 0 FP64 in
    draw order (the reference), 1 FP64 last draw first, 2 fixed point."""
    lib().og_set_deposit_order(int(mode))


def rounding_envelope(g: OgGrid, samples_col, idx_bs, deposits=False, **kw):
    """The oracle's estimate in its default summation orders and in others:
    the FFT radices reversed, radix-2 passes, the real part of each inverse
    transform kept instead of |.| of the complex result (whose imaginary part
    is rounding noise), and (deposits=True) each resample's deposits summed
    last draw first and rounded to the fixed-point quantum -- how far rounding
    alone moves the reference's result on this input.  Returns (default
    result, [results under the other orders])."""
    r0 = gradepro(g, samples_col, idx_bs=idx_bs, **kw)
    others = []
    try:
        for mode in (1, 2):
            set_fft_order(mode)
            others.append(gradepro(g, samples_col, idx_bs=idx_bs, **kw))
        set_fft_order(0)
        lib().og_set_real_output(1)
        others.append(gradepro(g, samples_col, idx_bs=idx_bs, **kw))
        lib().og_set_real_output(0)
        if deposits:
            for mode in (1, 2):
                lib().og_set_deposit_order(mode)
                others.append(gradepro(g, samples_col, idx_bs=idx_bs, **kw))
    finally:
        set_fft_order(0)
        lib().og_set_deposit_order(0)
        lib().og_set_real_output(0)
    return r0, others


def dirac_sequences(g: OgGrid, samples_col: np.ndarray, idx_bs: np.ndarray, with_squares=True, threads=1):
    """samples_col (n, ndims); idx_bs (B, n) 0-based.  Returns arrays
    [B][G] in Julia column-major linear order."""
    s = np.ascontiguousarray(samples_col, dtype=np.float64)
    idx = np.ascontiguousarray(idx_bs, dtype=np.int32)
    B, n = idx.shape
    G = int(np.prod(grid_shape(g)))
    sums = np.zeros(B * G)
    sqs = np.zeros(B * G) if with_squares else None
    rc = lib().og_dirac_sequences(C.byref(g), s.ravel(), n, idx.ravel(), B, sums,
                                  sqs.ctypes.data if sqs is not None else None, threads)
    return rc, sums.reshape(B, G), (sqs.reshape(B, G) if sqs is not None else None)


def rot(g: OgGrid, samples_col, rule=0):
    """og_rot_estimate: (rc, density [G] column-major, propagation time [N])."""
    s = np.ascontiguousarray(samples_col, dtype=np.float64)
    G = int(np.prod(grid_shape(g)))
    dens = np.zeros(G)
    t = np.zeros(g.ndims)
    rc = lib().og_rot_estimate(C.byref(g), s.ravel(), s.shape[0], int(rule), dens, t)
    return rc, dens, t


@dataclass
class OracleResult:
    density: np.ndarray
    chosen: np.ndarray
    vmr: np.ndarray | None
    means: np.ndarray | None
    info: OgInfo
    rc: int


def gradepro(g: OgGrid, samples_col, idx_bs=None, seed=0, n_bootstraps=100, n_steps=250,
             bw_step=float("nan"), bw_final=None, alpha_s=float("nan"), alpha_os=float("nan"),
             eps=float("nan"), threads=1, tables=False, max_rungs=0) -> OracleResult:
    s = np.ascontiguousarray(samples_col, dtype=np.float64)
    n = s.shape[0]
    p = OgParams()
    lib().og_params_default(C.byref(p))
    p.n_bootstraps = n_bootstraps
    p.n_steps = n_steps
    p.bw_step = bw_step
    if bw_final is not None:
        bf = np.atleast_1d(np.asarray(bw_final, dtype=np.float64))
        for i in range(g.ndims):
            p.bw_final[i] = float(bf[i] if len(bf) > 1 else bf[0])
    p.alpha_s, p.alpha_os, p.eps = alpha_s, alpha_os, eps
    p.threads = threads
    G = int(np.prod(grid_shape(g)))
    density = np.zeros(G)
    chosen = np.zeros(G, dtype=np.int16)
    info = OgInfo()
    idx_ptr = None
    idx_arr = None
    if idx_bs is not None:
        idx_arr = np.ascontiguousarray(idx_bs, dtype=np.int32)
        idx_ptr = idx_arr.ctypes.data
    # number of rungs for table allocation (only the first max_rungs rungs
    # are run and tabulated when max_rungs > 0)
    S_guess = (n_steps + 1) if n_steps > 0 else 4096
    if max_rungs > 0:
        S_guess = min(S_guess, max_rungs)
    vmr = np.zeros(S_guess * G) if tables else None
    means = np.zeros(S_guess * G) if tables else None
    rc = lib().og_gradepro_estimate(C.byref(g), s.ravel(), n, idx_ptr, seed, C.byref(p), density, chosen,
                                    vmr.ctypes.data if tables else None,
                                    means.ctypes.data if tables else None, max_rungs, C.byref(info))
    S = info.n_rungs
    if tables:
        S_tab = min(S, S_guess)
        vmr = vmr[: S_tab * G].reshape(S_tab, G)
        means = means[: S_tab * G].reshape(S_tab, G)
    return OracleResult(density, chosen, vmr, means, info, rc)
