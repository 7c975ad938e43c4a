"""Admission and failure isolation of the batched and serving forms.

The memory model is get_necessary_memory / check_memory
(GradeProEstimator.jl:1091-1112) extended to a resident set: a request's
footprint is computed on the host (pkde_gradepro_request_footprint); the
server admits requests while the resident plans and chunk buffers plus what
the request still needs fit its budget, rejects a request whose footprint
alone exceeds the budget with PKDE_OUT_OF_MEMORY (an ArgumentError, as the
reference's "Not enough memory" is) and completes the others.  A request that
fails in a batch or in the server does not change the others' results.
"""
import sys

import numpy as np
import pytest

import gpu_common as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pk():
    import parallelkde_rocm as pk
    pk.context()
    return pk


def _plain(pk, x, grid, **kw):
    est = pk.initialize_estimation(x, grid=grid)
    pk.estimate_density(est, "gradepro", **kw)
    return pk.get_density(est).copy(), est.last_result.chosen_rung.copy()


def _small_requests(pk):
    g1 = pk.initialize_grid([(-5.0, 5.0, 512)])
    g2 = pk.initialize_grid([(-6.0, 6.0, 64)] * 2)
    return [(W.gaussian_1d(4000, 81), g1, dict(n_bootstraps=20, n_steps=60, seed=11)),
            (W.mixture_2d(6000, 82), g2, dict(n_bootstraps=20, n_steps=60, seed=12)),
            (W.gaussian_1d(3000, 83), g1, dict(n_bootstraps=20, n_steps=60, seed=13))]


def test_request_footprint_is_host_side_and_monotone(pk):
    """The footprint needs no plan: it grows with the samples and the grid,
    and a 128^3 request needs gigabytes while a 512-point one needs
    megabytes."""
    g1 = pk.initialize_grid([(-5.0, 5.0, 512)])
    g3 = pk.initialize_grid([(-4.0, 4.0, 128)] * 3)
    x1 = W.gaussian_1d(1000, 1)
    f1 = pk.request_footprint(x1, g1)
    f1b = pk.request_footprint(W.gaussian_1d(100000, 1), g1)
    f3 = pk.request_footprint(W.correlated_3d(20000, 2), g3)
    say(f"  footprints: 1-D 512 n=1e3 {f1 / 2**20:.1f} MiB, n=1e5 {f1b / 2**20:.1f} MiB, "
        f"128^3 {f3 / 2**30:.2f} GiB")
    assert 0 < f1 < f1b < f3
    assert f1 < 64 << 20 and f3 > 8 << 30


def test_server_rejects_over_budget_request_others_complete(pk):
    """A request larger than the server's whole budget completes with
    PKDE_OUT_OF_MEMORY; the requests around it complete with the plain
    path's results; the resident bytes stay within the budget."""
    small = _small_requests(pk)
    budget = 2 * max(pk.request_footprint(x, g, **kw) for x, g, kw in small) + (64 << 20)
    g3 = pk.initialize_grid([(-4.0, 4.0, 128)] * 3)
    big = (W.correlated_3d(20000, 84), g3, dict(n_bootstraps=20, n_steps=60, seed=14))
    assert pk.request_footprint(big[0], big[1], **big[2]) > budget
    srv = pk.Server(max_inflight=2, memory_budget=budget)
    try:
        tickets = [srv.submit(x, g, **kw) for x, g, kw in small[:2]]
        t_big = srv.submit(*big[:2], **big[2])
        tickets.append(srv.submit(*small[2][:2], **small[2][2]))
        outs = [srv.wait(t) for t in tickets]
        with pytest.raises(pk.ArgumentError, match="memory"):
            srv.wait(t_big)
        st = srv.stats()
    finally:
        srv.close()
    say(f"  server stats {st}")
    assert st["rejected"] == 1 and st["completed"] == 3
    assert 0 < st["bytes_in_use"] <= budget
    for (x, g, kw), (d, c) in zip(small, outs):
        dp, cp = _plain(pk, x, g, **kw)
        assert np.array_equal(d, dp) and np.array_equal(c, cp)


def test_plain_estimate_over_device_memory_is_refused_before_allocating(pk):
    """check_memory (GradeProEstimator.jl:1098-1112): an estimate whose
    footprint exceeds the device's free memory (a 1024 x 1024 x 256 grid:
    its 201 binning fields alone are ~430 GB) raises the reference's
    ArgumentError("Not enough memory available. ...") before any device
    allocation, and the device's free memory and later estimates are
    untouched."""
    free0 = pk.device_free_memory()[0]
    grid = pk.initialize_grid([(-4.0, 4.0, 1024), (-4.0, 4.0, 1024), (-4.0, 4.0, 256)])
    est = pk.initialize_estimation(W.correlated_3d(1000, 91), grid=grid)
    with pytest.raises(pk.ArgumentError, match="Not enough memory available"):
        pk.estimate_density(est, "gradepro", n_bootstraps=100)
    est = None
    free1 = pk.device_free_memory()[0]
    say(f"  free device memory before {free0 >> 20} MiB, after {free1 >> 20} MiB")
    assert free1 >= free0 - (256 << 20)
    x, g, kw = _small_requests(pk)[0]
    d, c = _plain(pk, x, g, **kw)
    assert np.all(np.isfinite(d))


def test_server_failed_request_does_not_poison_others(pk):
    """A request with a sample outside its grid fails with an ArgumentError
    (check_data, DirectSpace.jl:479-498); the requests in flight with it
    complete with their plain results."""
    small = _small_requests(pk)
    bad_x = W.gaussian_1d(4000, 85)
    bad_x[17, 0] = 7.5  # outside [-5, 5]
    srv = pk.Server(max_inflight=3)
    try:
        t0 = srv.submit(*small[0][:2], **small[0][2])
        tb = srv.submit(bad_x, small[0][1], n_bootstraps=20, n_steps=60, seed=15)
        t1 = srv.submit(*small[1][:2], **small[1][2])
        t2 = srv.submit(*small[2][:2], **small[2][2])
        with pytest.raises(pk.ArgumentError, match="beyond grid"):
            srv.wait(tb)
        outs = [srv.wait(t) for t in (t0, t1, t2)]
    finally:
        srv.close()
    for (x, g, kw), (d, c) in zip(small, outs):
        dp, cp = _plain(pk, x, g, **kw)
        assert np.array_equal(d, dp) and np.array_equal(c, cp)


def test_batch_failed_request_does_not_poison_others(pk):
    """In one batched submission, an out-of-grid request and a request whose
    grid is refused (an axis above 4096 points) get their own status codes;
    the other requests equal the plain path bit for bit."""
    small = _small_requests(pk)
    bad_x = W.mixture_2d(5000, 86)
    bad_x[3, 1] = -9.0
    huge = pk.initialize_grid([(-5.0, 5.0, 5000)])
    items = [small[0], (bad_x, small[1][1], dict(small[1][2])), small[1],
             (W.gaussian_1d(2000, 87), huge, dict(n_bootstraps=20, n_steps=60)), small[2]]
    try:
        outs, status = pk.estimate_batch(items, with_chosen=True, return_status=True)
    finally:
        pk.release_batch_memory()
    say(f"  batch status {status}")
    assert status[1] == 2   # PKDE_OUT_OF_GRID
    assert status[3] == 6   # PKDE_UNSUPPORTED
    for k in (0, 2, 4):
        assert status[k] == 0
        x, g, kw = items[k]
        dp, cp = _plain(pk, x, g, **kw)
        assert np.array_equal(outs[k][0], dp) and np.array_equal(outs[k][1], cp)


def test_batch_many_seeds_share_plans(pk):
    """Requests that differ only in their bootstrap seed share one plan per
    stream (the seed is a launch argument): 48 requests of one 64^3 shape fit
    where 48 separate plans would not, and each equals its plain estimate."""
    g = pk.initialize_grid([(-6.0, 6.0, 64)] * 3)
    x = W.correlated_3d(20000, 88, lo=-6.0, hi=6.0)
    items = [(x, g, dict(n_bootstraps=40, n_steps=80, seed=100 + k)) for k in range(48)]
    try:
        outs = pk.estimate_batch(items, with_chosen=True)
    finally:
        pk.release_batch_memory()
    for k in (0, 17, 47):
        dp, cp = _plain(pk, x, g, **items[k][2])
        assert np.array_equal(outs[k][0], dp) and np.array_equal(outs[k][1], cp)
    assert not np.array_equal(outs[0][0], outs[1][0])
