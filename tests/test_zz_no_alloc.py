"""The allocation-free entry point (pkde_gradepro_estimate_1d_noalloc)
against the plan-based device entry point (pkde_gradepro_estimate_device) it
is a variant of, through the C ABI: bit-identical densities and chosen rungs
on two calls in a row, the second with a smaller input; limits refused with
an error; no device memory taken by repeated calls."""
import ctypes as C
import sys

import numpy as np
import pytest

import gpu_common as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import parallelkde_rocm as pk
    from parallelkde_rocm import _lib
    hip = C.CDLL("libamdhip64.so.7")  # the runtime the library itself uses
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return pk, _lib, _lib.load(), hip


class _Dev:
    """A device buffer of `nbytes` from the HIP runtime the library uses."""
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(pk, _lib, L, hip, pts, x, seed, steps, d_x, d_dens, d_ch):
    grid = pk.initialize_grid([(-5.0, 5.0, pts)])
    params = pk._params(1, n_bootstraps=100, n_steps=steps, seed=seed)
    d_x.put(x)
    rc = L.pkde_gradepro_estimate_1d_noalloc(pk.context().handle, pts, grid.lo[0], grid.hi[0],
                                             grid.h[0], C.byref(params), d_x.p, x.shape[0],
                                             d_dens.p, d_ch.p, None)
    assert rc == 0, _lib.last_error()
    assert hip.hipDeviceSynchronize() == 0
    return d_dens.get(pts, np.float64), d_ch.get(pts, np.int16)


def _plan(pk, _lib, L, hip, pts, x, seed, steps, d_x, d_dens, d_ch):
    grid = pk.initialize_grid([(-5.0, 5.0, pts)])
    plan = pk._Plan(grid, pk._params(1, n_bootstraps=100, n_steps=steps, seed=seed))
    try:
        d_x.put(x)
        rc = L.pkde_gradepro_estimate_device(plan.handle, d_x.p, x.shape[0], None, d_dens.p,
                                             d_ch.p, None)
        assert rc == 0, _lib.last_error()
        assert L.pkde_gradepro_check(plan.handle) == 0
        return d_dens.get(pts, np.float64), d_ch.get(pts, np.int16)
    finally:
        plan.close()


def test_noalloc_equals_plan_entry_point(env):
    pk, _lib, L, hip = env
    d_x, d_dens, d_ch = _Dev(hip, 131072 * 8), _Dev(hip, 1024 * 8), _Dev(hip, 1024 * 2)
    try:
        # two calls in a row, the second with a smaller input (fewer samples,
        # a smaller grid, a shorter ladder): the array keeps the first call's
        # contents, which the second must not see
        for pts, n, seed, steps in ((1024, 100000, 21, 250), (256, 3000, 22, 120)):
            x = W.gaussian_1d(n, seed)
            a = _noalloc(pk, _lib, L, hip, pts, x, seed, steps, d_x, d_dens, d_ch)
            b = _plan(pk, _lib, L, hip, pts, x, seed, steps, d_x, d_dens, d_ch)
            say(f"  {pts} points, n={n}: equal {np.array_equal(a[0], b[0])}")
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            assert np.all(np.isfinite(a[0]))
    finally:
        for d in (d_x, d_dens, d_ch):
            d.free()


def test_noalloc_takes_no_device_memory_and_refuses_limits(env):
    pk, _lib, L, hip = env
    d_x, d_dens, d_ch = _Dev(hip, 131072 * 8), _Dev(hip, 1024 * 8), _Dev(hip, 1024 * 2)
    try:
        x = W.gaussian_1d(20000, 31)
        _noalloc(pk, _lib, L, hip, 512, x, 31, 250, d_x, d_dens, d_ch)
        free0, tot = C.c_uint64(), C.c_uint64()
        assert L.pkde_device_free_memory(pk.context().handle, C.byref(free0), C.byref(tot)) == 0
        for k in range(3):
            _noalloc(pk, _lib, L, hip, 512, x, 32 + k, 250, d_x, d_dens, d_ch)
        free1 = C.c_uint64()
        assert L.pkde_device_free_memory(pk.context().handle, C.byref(free1), C.byref(tot)) == 0
        assert free1.value == free0.value
        grid = pk.initialize_grid([(-5.0, 5.0, 512)])
        params = pk._params(1, n_bootstraps=100, n_steps=250)
        ctx = pk.context().handle
        # more samples than the array takes, a grid outside the range, a
        # ladder too long for the array
        assert L.pkde_gradepro_estimate_1d_noalloc(ctx, 512, grid.lo[0], grid.hi[0], grid.h[0],
                                                   C.byref(params), d_x.p, 200000, d_dens.p,
                                                   d_ch.p, None) == 1
        assert L.pkde_gradepro_estimate_1d_noalloc(ctx, 2048, grid.lo[0], grid.hi[0], grid.h[0],
                                                   C.byref(params), d_x.p, 1000, d_dens.p,
                                                   d_ch.p, None) == 6
        long_params = pk._params(1, n_bootstraps=100, n_steps=4000)
        assert L.pkde_gradepro_estimate_1d_noalloc(ctx, 1024, grid.lo[0], grid.hi[0], grid.h[0],
                                                   C.byref(long_params), d_x.p, 100000, d_dens.p,
                                                   d_ch.p, None) == 4
        assert "scratch" in _lib.last_error()
    finally:
        for d in (d_x, d_dens, d_ch):
            d.free()
