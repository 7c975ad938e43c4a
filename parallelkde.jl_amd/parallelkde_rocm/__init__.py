"""Python mirror of ParallelKDE.jl's user API for the `:rocm` device.

Names, argument meaning and error behaviour follow the reference
(src/ParallelKDE.jl:102-226, src/Grids/Grids.jl, src/KDEs/KDEs.jl,
src/Devices.jl, src/DensityEstimators/DensityEstimators.jl); `estimate_density!`
is spelled `estimate_density` (alias `estimate_density_`).  Data follow the
ParallelKDEpy convention, an array of shape (n_samples, n_dims), which has the
same memory layout as Julia's N x n matrix.  Densities come back as arrays of
shape (n_1, ..., n_N) in Fortran order, i.e. Julia's Array{T,N} indexing.

Every estimate runs on the GPU through libparallelkde_hip; there is no CPU
fallback in this package (the reference's own CPU path is the Julia code).
"""
from __future__ import annotations

import math
import sys
import threading
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import OK, INVALID_ARGUMENT, OUT_OF_GRID, NO_DEVICE, OUT_OF_MEMORY, UNSUPPORTED

__all__ = [
    "ArgumentError", "DeviceUnavailableError", "Grid", "KDE", "DensityEstimation",
    "initialize_grid", "find_grid", "initialize_kde", "initialize_estimation",
    "estimate_density", "estimate_density_", "get_density", "has_grid", "get_grid", "autotune",
    "spacings", "bounds", "low_bounds", "high_bounds", "initial_bandwidth", "fftgrid",
    "AVAILABLE_DEVICES", "DEVICE_IMPLEMENTATIONS", "DEFAULT_IMPLEMENTATIONS",
    "estimator_lookup", "GradeProResult", "RoTResult", "estimate_batch", "Server", "context",
    "request_footprint", "release_batch_memory",
]


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class DeviceUnavailableError(RuntimeError):
    """No usable ROCm device.  The reference warns and falls back to :cpu
    (ParallelKDE.jl:123-127); this package has no CPU path and raises."""


# ---------------------------------------------------------------- devices
# Devices.jl:54-82 with the new :rocm entry.
AVAILABLE_DEVICES = {"rocm": "IsROCm"}
DEVICE_IMPLEMENTATIONS = {"rocm": {"rocm"}}
DEFAULT_IMPLEMENTATIONS = {"rocm": "rocm"}


def _device(device) -> str:
    d = str(device).lstrip(":")
    if d not in AVAILABLE_DEVICES:
        raise ArgumentError(f"Invalid device: {device}")
    return d


def ensure_valid_implementation(device, implementation):
    """Devices.jl:125-133."""
    d = _device(device)
    if str(implementation).lstrip(":") not in DEVICE_IMPLEMENTATIONS[d]:
        raise ArgumentError(f"Invalid implementation {implementation} for device {device}")
    return True


def _raise(rc: int, what: str):
    if rc == OK:
        return
    msg = _lib.last_error() or what
    if rc == NO_DEVICE:
        raise DeviceUnavailableError(msg)
    if rc in (INVALID_ARGUMENT, OUT_OF_GRID, OUT_OF_MEMORY, UNSUPPORTED):
        raise ArgumentError(msg)
    raise RuntimeError(f"{what}: {msg} (status {rc})")


class _Context:
    def __init__(self, device: int = 0):
        L = _lib.load()
        h = _lib.C.c_void_p()
        _raise(L.pkde_context_create(device, _lib.C.byref(h)), "context")
        self.handle = h
        self.lock = threading.Lock()

    def __del__(self):
        # at interpreter shutdown finalizers run in no fixed order; the
        # process exit releases device memory, so native calls are skipped
        if sys.is_finalizing():
            return
        try:
            if self.handle:
                _lib.load().pkde_context_destroy(self.handle)
        except Exception:
            pass


_ctx = None
_ctx_lock = threading.Lock()


def context() -> _Context:
    global _ctx
    with _ctx_lock:
        if _ctx is None:
            _ctx = _Context(0)
        return _ctx


def device_free_memory():
    """(free, total) bytes of the device's memory: get_available_memory
    (Devices.jl:103-108) for the :rocm device."""
    free, total = _lib.C.c_uint64(), _lib.C.c_uint64()
    _raise(_lib.load().pkde_device_free_memory(context().handle, _lib.C.byref(free),
                                                _lib.C.byref(total)), "free memory")
    return int(free.value), int(total.value)


# ------------------------------------------------------------------ grids
@dataclass
class Grid:
    """Regular Cartesian grid (Grids.jl:67-87): per axis lo, hi, spacing, n."""
    lo: tuple
    hi: tuple
    h: tuple
    n: tuple
    device: str = "rocm"

    @property
    def shape(self):
        return tuple(int(v) for v in self.n)

    def size(self):
        return self.shape

    @property
    def ndims(self):
        return len(self.n)

    def get_coordinates(self):
        """N x n_1 x ... x n_N coordinates (Grids.jl:157-169)."""
        axes = [self.lo[i] + self.h[i] * np.arange(self.n[i]) for i in range(self.ndims)]
        for i in range(self.ndims):
            axes[i][-1] = self.hi[i]
        mesh = np.meshgrid(*axes, indexing="ij")
        return np.stack(mesh)

    def to_meshgrid(self):
        return list(self.get_coordinates())


def _range_spec(r):
    """(lo, hi, n, h) from a (lo, hi, n) tuple or an evenly spaced vector."""
    if isinstance(r, (tuple, list)) and len(r) == 3 and not isinstance(r[2], float):
        lo, hi, n = float(r[0]), float(r[1]), int(r[2])
        if n < 2:
            raise ArgumentError("a grid axis needs at least two points")
        return lo, hi, n, (hi - lo) / (n - 1)
    a = np.asarray(r, dtype=np.float64)
    if a.ndim != 1 or len(a) < 2:
        raise ArgumentError("Elements of ranges must implement 'step' method.")
    d = np.diff(a)
    if not np.allclose(d, d[0], rtol=1e-9, atol=0):
        raise ArgumentError("Elements of ranges must implement 'step' method.")
    n = len(a)
    return float(a[0]), float(a[-1]), n, (float(a[-1]) - float(a[0])) / (n - 1)


def initialize_grid(ranges, device="rocm", **_):
    """Grids.jl:38-50."""
    d = _device(device)
    if isinstance(ranges, tuple) and len(ranges) == 3 and np.isscalar(ranges[0]):
        ranges = [ranges]
    specs = [_range_spec(r) for r in ranges]
    return Grid(tuple(s[0] for s in specs), tuple(s[1] for s in specs),
                tuple(s[3] for s in specs), tuple(s[2] for s in specs), d)


#: Default points per axis of find_grid, as in the reference (Grids.jl:282; its
#: docstring says 300, the code uses 150).  150 = 2 x 3 x 5^2 runs on the
#: mixed-radix FFT path.
FIND_GRID_DEFAULT_POINTS = 150


def _device_bounds(data, grid_padding):
    """Padded bounds of samples already in device memory (a float64 tensor of
    shape (n, ndims) on the GPU): pkde_find_grid reduces the extrema there.
    torch carries its own HIP runtime: initialise torch's GPU state before
    this package creates its context, so that both use one runtime."""
    x = data if data.dim() == 2 else data.reshape(-1, 1)
    if x.dtype != _torch_float64() or not x.is_contiguous():
        raise ArgumentError("device samples must be a contiguous float64 tensor of shape (n, ndims)")
    n, n_dims = int(x.shape[0]), int(x.shape[1])
    lo = (_lib.C.c_double * n_dims)()
    hi = (_lib.C.c_double * n_dims)()
    pad = None
    if grid_padding is not None:
        pad = (_lib.C.c_double * n_dims)(*[float(p) for p in grid_padding])
    _raise(_lib.load().pkde_find_grid(context().handle, n_dims, _lib.C.c_void_p(x.data_ptr()), n,
                                       pad, lo, hi), "find_grid")
    return [(lo[i], hi[i]) for i in range(n_dims)]


def _torch_float64():
    import torch
    return torch.float64


def _on_device(data) -> bool:
    return bool(getattr(data, "is_cuda", False))


def find_grid(data, grid_bounds=None, grid_dims=None, grid_steps=None, grid_padding=None,
              device="rocm"):
    """Grids.jl:252-300: data extrema padded by 10% of the range per axis.
    Samples already in device memory (a torch tensor on the GPU) are reduced
    on the device (pkde_find_grid); host samples on the host."""
    if grid_bounds is None and _on_device(data):
        gb = _device_bounds(data, grid_padding)
        n_dims = len(gb)
    else:
        x = _as_samples(data)
        n_dims = x.shape[1]
        if grid_bounds is None:
            grid_bounds = [(float(x[:, i].min()), float(x[:, i].max())) for i in range(n_dims)]
        if grid_padding is None:
            grid_padding = [0.1 * abs(b[1] - b[0]) for b in grid_bounds]
        gb = [(b[0] - p, b[1] + p) for b, p in zip(grid_bounds, grid_padding)]
    if grid_steps is None and grid_dims is None:
        grid_dims = (FIND_GRID_DEFAULT_POINTS,) * n_dims
    if grid_dims is not None:
        grid_dims = tuple(grid_dims) if np.iterable(grid_dims) else (int(grid_dims),) * n_dims
        ranges = [(gb[i][0], gb[i][1], int(grid_dims[i])) for i in range(n_dims)]
    else:
        steps = grid_steps if np.iterable(grid_steps) else (grid_steps,) * n_dims
        ranges = []
        for i in range(n_dims):
            n = int(math.floor((gb[i][1] - gb[i][0]) / steps[i] + 1e-10)) + 1
            ranges.append(gb[i][0] + steps[i] * np.arange(n))
    return initialize_grid(ranges, device=device)


def spacings(grid: Grid):
    return np.array(grid.h)


def bounds(grid: Grid):
    return np.array([grid.lo, grid.hi])


def low_bounds(grid: Grid):
    return np.array(grid.lo)


def high_bounds(grid: Grid):
    return np.array(grid.hi)


def initial_bandwidth(grid: Grid):
    """Grids.jl:217: half the spacing."""
    return np.array(grid.h) / 2


def fftgrid(grid: Grid):
    """Grids.jl:225-231: 2pi * fftfreq(n, 1/h) per axis."""
    out = []
    for i in range(grid.ndims):
        n, h = grid.n[i], grid.h[i]
        mult = ((1.0 / h) / n) * 6.283185307179586
        j = np.arange(n)
        out.append(np.where(j < (n + 1) // 2, j, j - n) * mult)
    return out


# ------------------------------------------------------------------- KDEs
def _as_samples(data) -> np.ndarray:
    if isinstance(data, (list, tuple)) and len(data) and np.ndim(data[0]) == 1:
        data = np.asarray(data, dtype=np.float64)
    x = np.asarray(data, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    if x.ndim != 2:
        raise ArgumentError("Data must be a matrix or vector.")
    return np.ascontiguousarray(x)


@dataclass
class KDE:
    """KDEs.jl:33-36: samples and a NaN-initialised density."""
    data: np.ndarray
    density: np.ndarray
    device: str = "rocm"

    def get_nsamples(self):
        return self.data.shape[0]


def initialize_kde(data, dims, device="rocm"):
    """KDEs.jl:62-111."""
    d = _device(device)
    x = _as_samples(data)
    dims = tuple(int(v) for v in (dims if np.iterable(dims) else (dims,)))
    if x.shape[1] != len(dims):
        raise ArgumentError(f"Data must have {len(dims)} dimensions")
    return KDE(x, np.full(dims, np.nan, order="F"), d)


@dataclass
class DensityEstimation:
    """ParallelKDE.jl:75-78."""
    kde: KDE
    grid: Grid | None
    _plans: dict = field(default_factory=dict, repr=False)
    last_result: "GradeProResult | None" = None

    def estimate_density(self, method="gradepro", **kwargs):
        estimate_density(self, method, **kwargs)
        return self

    def get_density(self, normalize=False, dx=None):
        return get_density(self, normalize=normalize, dx=dx)

    def __del__(self):
        # at interpreter shutdown finalizers run in no fixed order; the
        # process exit releases device memory, so native calls are skipped
        if sys.is_finalizing():
            return
        try:
            for p in self._plans.values():
                p.close()
        except Exception:
            pass


def initialize_estimation(data, grid=False, grid_ranges=None, dims=None, grid_bounds=None,
                          grid_padding=None, device="rocm"):
    """ParallelKDE.jl:102-158."""
    x = _as_samples(data)
    d = _device(device)
    if isinstance(grid, bool):
        if grid:
            if grid_ranges is not None:
                grid = initialize_grid(grid_ranges, device=d)
            else:
                grid = find_grid(x, grid_bounds=grid_bounds, grid_dims=dims,
                                 grid_padding=grid_padding, device=d)
        else:
            grid = None
    elif isinstance(grid, Grid):
        if grid.device != d:
            raise ArgumentError("Grid must be of the same device as the KDE.")
    else:
        raise ArgumentError("grid must be a Bool or a Grid")
    if grid is None and dims is None:
        raise ArgumentError("Grid must be required if dims is not specified.")
    if grid is not None:
        dims = grid.shape
    return DensityEstimation(initialize_kde(x, dims, device=d), grid)


def has_grid(est: DensityEstimation) -> bool:
    return est.grid is not None


def get_grid(est: DensityEstimation):
    return est.grid


def get_density(est: DensityEstimation, normalize=False, dx=None):
    """ParallelKDE.jl:189-202.  Returns the stored density (normalised in
    place when asked, as the reference does)."""
    density = est.kde.density
    if normalize and has_grid(est):
        density /= density.sum() * np.prod(spacings(est.grid))
    elif normalize and dx is not None:
        density /= density.sum() * dx
    elif normalize:
        raise ArgumentError("Normalization requires a grid or dx to be specified.")
    return density


# -------------------------------------------------------------- estimator
@dataclass
class GradeProResult:
    density: np.ndarray
    chosen_rung: np.ndarray
    n_rungs: int
    steps_low: int
    steps_over: int
    eps: float
    bw_final: np.ndarray
    ms_total: float
    ms_ladder: float
    algorithmic_bytes: float


def _params(ndims, n_bootstraps=100, bw_final=None, bw_step=None, n_steps=None, alpha_s=None,
            alpha_os=None, eps=None, seed=0x5EED, precision="fp64", deterministic=True):
    p = _lib.Params()
    _lib.load().pkde_gradepro_params_default(_lib.C.byref(p))
    if int(n_bootstraps) < 0:
        raise ArgumentError("Number of boostraps must be positive.")
    p.n_bootstraps = int(n_bootstraps)
    if n_steps is not None:
        p.n_steps = int(n_steps)
    elif bw_step is not None:
        p.n_steps = 0
    if bw_step is not None:
        p.bw_step = float(bw_step)
    if bw_final is not None:
        bf = np.atleast_1d(np.asarray(bw_final, dtype=np.float64))
        for i in range(ndims):
            p.bw_final[i] = float(bf[i] if len(bf) > 1 else bf[0])
    if alpha_s is not None:
        p.alpha_s = float(alpha_s)
    if alpha_os is not None:
        p.alpha_os = float(alpha_os)
    if eps is not None:
        p.eps = float(eps)
    p.rng_seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    p.deterministic_binning = 1 if deterministic else 0
    if precision not in ("fp32", "fp64"):
        raise ArgumentError("precision must be 'fp32' or 'fp64'")
    p.precision = 1 if precision == "fp64" else 0
    return p


class _Plan:
    def __init__(self, grid: Grid, params: _lib.Params):
        L = _lib.load()
        self.ctx = context()
        N = grid.ndims
        n = (_lib.C.c_int64 * N)(*grid.n)
        lo = (_lib.C.c_double * N)(*grid.lo)
        hi = (_lib.C.c_double * N)(*grid.hi)
        h = (_lib.C.c_double * N)(*grid.h)
        self.handle = _lib.C.c_void_p()
        _raise(L.pkde_gradepro_plan_create(self.ctx.handle, N, n, lo, hi, h, _lib.C.byref(params),
                                           _lib.C.byref(self.handle)), "plan")
        self.grid = grid
        self.B = params.n_bootstraps

    def estimate(self, x: np.ndarray, bootstrap_idx=None):
        L = _lib.load()
        G = int(np.prod(self.grid.shape))
        dens = np.empty(G)
        chosen = np.empty(G, dtype=np.int16)
        st = _lib.Stats()
        idx = None
        if bootstrap_idx is not None:
            idx = np.ascontiguousarray(bootstrap_idx, dtype=np.int32)
            if idx.shape != (self.B, x.shape[0]):
                raise ArgumentError("bootstrap_idx must have shape (n_bootstraps, n_samples)")
        rc = L.pkde_gradepro_estimate(self.handle, x.ctypes.data, x.shape[0], _lib.ptr(idx),
                                      dens.ctypes.data, chosen.ctypes.data, _lib.C.byref(st))
        _raise(rc, "estimate")
        shape = self.grid.shape
        return GradeProResult(dens.reshape(shape, order="F"), chosen.reshape(shape, order="F"),
                              st.n_rungs, st.steps_low, st.steps_over, st.eps,
                              np.array(st.bw_final[: self.grid.ndims]), st.ms_total, st.ms_ladder,
                              st.algorithmic_bytes)

    def autotune(self, reps=3):
        """Ladder launch-shape autotuner (pkde_gradepro_autotune): returns the
        ladder milliseconds of the chosen shape and of the default one."""
        best = _lib.C.c_double()
        default = _lib.C.c_double()
        _raise(_lib.load().pkde_gradepro_autotune(self.handle, int(reps), _lib.C.byref(best),
                                                  _lib.C.byref(default)), "autotune")
        return best.value, default.value

    def fft_forward(self, fields):
        """Forward DFT of real fields with the estimate's own kernels
        (pkde_fft_forward): fields (count, G) with grid axis 1 fastest ->
        complex (count, n1 // 2 + 1, G // n1), [field][k1][q]."""
        f = np.ascontiguousarray(np.asarray(fields, dtype=np.float64).reshape(len(fields), -1))
        n1 = self.grid.shape[0]
        out = np.empty((f.shape[0], n1 // 2 + 1, f.shape[1] // n1), dtype=np.complex128)
        _raise(_lib.load().pkde_fft_forward(self.handle, f.shape[0], f.ctypes.data,
                                            out.ctypes.data), "forward transform")
        return out

    def debug_vmr(self, n_rungs):
        G = int(np.prod(self.grid.shape))
        out = np.empty(n_rungs * G)
        _raise(_lib.load().pkde_gradepro_debug_vmr(self.handle, out.ctypes.data, out.size), "vmr")
        return out.reshape(n_rungs, G)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().pkde_gradepro_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        # at interpreter shutdown finalizers run in no fixed order; the
        # process exit releases device memory, so native calls are skipped
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class GradeProEstimator:
    """Registry entry for :gradepro (GradeProEstimator.jl:695)."""
    KWARGS = ("n_bootstraps", "bw_final", "bw_step", "n_steps", "alpha_s", "alpha_os", "eps",
              "seed", "precision", "deterministic")


class RoTEstimator:
    """Registry entry for :rot (RoTEstimator.jl:111).  `rule_of_thumb` is a
    keyword of initialize_estimator there (default :scott, :120); here it is
    a keyword of estimate_density."""
    KWARGS = ("rule_of_thumb", "precision")
    RULES = {"scott": 0, "silverman": 1}


# DensityEstimators.jl:31-45
estimator_lookup = {"gradepro": GradeProEstimator, "rot": RoTEstimator}


@dataclass
class RoTResult:
    density: np.ndarray
    time_propagated: np.ndarray  # rule-of-thumb bandwidth minus h/2, per axis
    ms_total: float


class _RoTPlan:
    def __init__(self, grid: Grid, rule: int, precision: int):
        L = _lib.load()
        self.ctx = context()
        N = grid.ndims
        n = (_lib.C.c_int64 * N)(*grid.n)
        lo = (_lib.C.c_double * N)(*grid.lo)
        hi = (_lib.C.c_double * N)(*grid.hi)
        h = (_lib.C.c_double * N)(*grid.h)
        self.handle = _lib.C.c_void_p()
        _raise(L.pkde_rot_plan_create(self.ctx.handle, N, n, lo, hi, h, rule, precision,
                                      _lib.C.byref(self.handle)), "plan")
        self.grid = grid

    def estimate(self, x: np.ndarray) -> RoTResult:
        L = _lib.load()
        G = int(np.prod(self.grid.shape))
        dens = np.empty(G)
        st = _lib.Stats()
        rc = L.pkde_gradepro_estimate(self.handle, x.ctypes.data, x.shape[0], None,
                                      dens.ctypes.data, None, _lib.C.byref(st))
        _raise(rc, "estimate")
        return RoTResult(dens.reshape(self.grid.shape, order="F"),
                         np.array(st.bw_final[: self.grid.ndims]), st.ms_total)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().pkde_gradepro_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        # at interpreter shutdown finalizers run in no fixed order; the
        # process exit releases device memory, so native calls are skipped
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def _estimate_rot(est, rule_of_thumb="scott", precision="fp64"):
    """RoTEstimator.jl:116-265 on the GPU."""
    rule = str(rule_of_thumb).lstrip(":")
    if rule not in RoTEstimator.RULES:
        raise ArgumentError(f"Rule of thumb {rule} not implemented.")
    prec = {"fp64": 1, "fp32": 0}.get(precision)
    if prec is None:
        raise ArgumentError("precision must be 'fp64' or 'fp32'")
    pkey = ("rot", RoTEstimator.RULES[rule], prec)
    plan = est._plans.get(pkey)
    if plan is None:
        plan = _RoTPlan(est.grid, RoTEstimator.RULES[rule], prec)
        est._plans[pkey] = plan
    res = plan.estimate(est.kde.data)
    est.kde.density[...] = res.density
    est.last_result = res


def estimate_density(est: DensityEstimation, estimation_method="gradepro", method=None,
                     bootstrap_idx=None, **kwargs):
    """ParallelKDE.jl:211-226 -> DensityEstimators.jl:58-84 ->
    GradeProEstimator.jl:700-1072, on the GPU.  Extra keywords: `seed`
    (bootstrap RNG), `precision` ('fp32' | 'fp64'), `deterministic`,
    `bootstrap_idx` (explicit 0-based draws, shape (B, n))."""
    est.kde.density[...] = np.nan  # set_nan_density! (ParallelKDE.jl:216)
    key = str(estimation_method).lstrip(":")
    if key not in estimator_lookup:
        raise KeyError(key)
    device = est.kde.device
    if method is None:
        method = DEFAULT_IMPLEMENTATIONS[device]
    ensure_valid_implementation(device, method)
    if est.grid is None:
        raise ArgumentError(f"the :{key} estimator needs a grid")
    unknown = set(kwargs) - set(estimator_lookup[key].KWARGS)
    if unknown:
        raise TypeError(f"unexpected keyword arguments {sorted(unknown)}")
    if key == "rot":
        if bootstrap_idx is not None:
            raise TypeError("the :rot estimator takes no bootstrap_idx")
        _estimate_rot(est, **kwargs)
        return None
    if unknown:
        raise TypeError(f"unexpected keyword arguments {sorted(unknown)}")
    params = _params(est.grid.ndims, **kwargs)
    pkey = bytes(params)
    plan = est._plans.get(pkey)
    if plan is None:
        plan = _Plan(est.grid, params)
        est._plans[pkey] = plan
    res = plan.estimate(est.kde.data, bootstrap_idx)
    est.kde.density[...] = res.density
    est.last_result = res
    return None


estimate_density_ = estimate_density


def autotune(est: DensityEstimation, reps=3):
    """Tune the ladder launch shape of every :gradepro plan the estimation
    has run (run estimate_density first).  The choice is kept for plans of the
    same grid, B, ladder length and precision in this process's device
    context, the batched and serving forms included; densities are unchanged.
    Returns [(ladder ms chosen, ladder ms default), ...].  No counterpart in
    the reference."""
    out = []
    for plan in est._plans.values():
        if isinstance(plan, _Plan):
            out.append(plan.autotune(reps))
    if not out:
        raise ArgumentError("run estimate_density(est, :gradepro) before autotune")
    return out


# ------------------------------------------------------------ batched form
def _request(x, grid, params):
    r = _lib.Request()
    r.ndims = grid.ndims
    for a in range(grid.ndims):
        r.n_per_axis[a], r.lo[a], r.hi[a], r.h[a] = grid.n[a], grid.lo[a], grid.hi[a], grid.h[a]
    r.samples = x.ctypes.data
    r.n = x.shape[0]
    r.params = params
    return r


def request_footprint(data, grid: Grid, **kwargs):
    """Device bytes one estimate of `data` on `grid` needs in the batched and
    serving forms (its plan and the chunk buffer its stream lends it),
    computed on the host (pkde_gradepro_request_footprint)."""
    x = _as_samples(data)
    r = _request(x, grid, _params(grid.ndims, **kwargs))
    dummy = np.empty(1)
    r.density_out = dummy.ctypes.data
    out = _lib.C.c_uint64()
    _raise(_lib.load().pkde_gradepro_request_footprint(context().handle, _lib.C.byref(r),
                                                       _lib.C.byref(out)), "footprint")
    return int(out.value)


def release_batch_memory():
    """Free the plans and chunk buffers the batched form keeps resident in the
    device context between calls (pkde_context_release_batch)."""
    _raise(_lib.load().pkde_context_release_batch(context().handle), "release")


def estimate_batch(items, with_chosen=False, return_status=False, **common):
    """Many independent estimates in one submission.  items: iterable of
    (data, grid) or (data, grid, kwargs).  Returns a list of densities, or of
    (density, chosen rung) pairs with with_chosen=True."""
    L = _lib.load()
    ctx = context()
    reqs = (_lib.Request * len(items))()
    keep = []
    outs = []
    for i, it in enumerate(items):
        data, grid = it[0], it[1]
        kw = dict(common)
        if len(it) > 2:
            kw.update(it[2])
        x = _as_samples(data)
        p = _params(grid.ndims, **kw)
        r = reqs[i]
        r.ndims = grid.ndims
        for a in range(grid.ndims):
            r.n_per_axis[a], r.lo[a], r.hi[a], r.h[a] = grid.n[a], grid.lo[a], grid.hi[a], grid.h[a]
        r.samples = x.ctypes.data
        r.n = x.shape[0]
        r.params = p
        dens = np.empty(int(np.prod(grid.shape)))
        r.density_out = dens.ctypes.data
        chosen = np.empty(int(np.prod(grid.shape)), dtype=np.int16) if with_chosen else None
        r.chosen_out = chosen.ctypes.data if with_chosen else None
        keep.append(x)
        outs.append((dens, chosen, grid.shape))
    status = np.zeros(len(items), dtype=np.int32)
    rc = L.pkde_gradepro_estimate_batch(ctx.handle, len(items), reqs, status.ctypes.data)
    if return_status:
        # per-request status codes; a failed request leaves its output unset
        outs_r = [(d.reshape(shp, order="F"), None if c is None else c.reshape(shp, order="F"))
                  for d, c, shp in outs]
        return outs_r, [int(v) for v in status]
    for s in status:
        _raise(int(s), "batch request")
    _raise(rc, "batch")
    if with_chosen:
        return [(d.reshape(shp, order="F"), c.reshape(shp, order="F")) for d, c, shp in outs]
    return [d.reshape(shp, order="F") for d, _, shp in outs]


class Server:
    """Continuous-batching server over the C ABI (pkde_server_*)."""

    def __init__(self, max_inflight=4, memory_budget=0):
        L = _lib.load()
        self.ctx = context()
        cfg = _lib.ServerCfg(int(max_inflight), int(memory_budget))
        self.handle = _lib.C.c_void_p()
        _raise(L.pkde_server_create(self.ctx.handle, _lib.C.byref(cfg), _lib.C.byref(self.handle)),
               "server")
        self._pending = {}

    def submit(self, data, grid: Grid, **kwargs):
        L = _lib.load()
        x = _as_samples(data)
        r = _lib.Request()
        r.ndims = grid.ndims
        for a in range(grid.ndims):
            r.n_per_axis[a], r.lo[a], r.hi[a], r.h[a] = grid.n[a], grid.lo[a], grid.hi[a], grid.h[a]
        r.samples = x.ctypes.data
        r.n = x.shape[0]
        r.params = _params(grid.ndims, **kwargs)
        dens = np.empty(int(np.prod(grid.shape)))
        chosen = np.empty(int(np.prod(grid.shape)), dtype=np.int16)
        r.density_out = dens.ctypes.data
        r.chosen_out = chosen.ctypes.data
        t = _lib.C.c_uint64()
        _raise(L.pkde_server_submit(self.handle, _lib.C.byref(r), _lib.C.byref(t)), "submit")
        self._pending[t.value] = (dens, chosen, grid.shape)
        return t.value

    def poll(self, ticket):
        done = _lib.C.c_int()
        rc = _lib.load().pkde_server_poll(self.handle, ticket, _lib.C.byref(done))
        if not done.value:
            return None
        dens, chosen, shp = self._pending.pop(ticket)
        _raise(rc, "request")
        return dens.reshape(shp, order="F"), chosen.reshape(shp, order="F")

    def wait(self, ticket):
        rc = _lib.load().pkde_server_wait(self.handle, ticket)
        dens, chosen, shp = self._pending.pop(ticket)
        _raise(rc, "request")
        return dens.reshape(shp, order="F"), chosen.reshape(shp, order="F")

    def stats(self):
        s = _lib.ServerStats()
        _raise(_lib.load().pkde_server_stats_get(self.handle, _lib.C.byref(s)), "stats")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().pkde_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        # at interpreter shutdown finalizers run in no fixed order; the
        # process exit releases device memory, so native calls are skipped
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
