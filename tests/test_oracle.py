"""Pins the FP64 CPU oracle to the reference's own known-answer tests and to
checks that fix the mathematics independently of the oracle's formulas.

Each test names the reference test it restates.  The reference ships no golden
files; the committed fixtures under tests/golden/ come from numpy and from the
Random123 publication (tests/golden/make_fixtures.py).
"""
import json
import math
import os

import numpy as np
import pytest

import gpu_common as W
import oracle_ctypes as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --------------------------------------------------------------------- RNG
def test_philox_published_kat():
    with open(os.path.join(GOLD, "philox_kat.json")) as f:
        kat = json.load(f)
    for ctr, key, expected in kat:
        assert O.philox(ctr, key) == expected


def test_bootstrap_indices_range_and_uniformity():
    # kde_tests.jl:17-20: size (n, B), integer, inside 1:n (0-based here)
    idx = O.bootstrap_indices(seed=11, n=1000, B=40)
    assert idx.shape == (40, 1000)
    assert idx.min() >= 0 and idx.max() < 1000
    counts = np.bincount(idx.ravel(), minlength=1000)
    # each index expected 40 times; chi-square well inside its bulk
    chi2 = ((counts - 40.0) ** 2 / 40.0).sum()
    assert 800 < chi2 < 1200
    assert not np.array_equal(idx[0], idx[1])


# --------------------------------------------------------------------- FFT
def test_fft_matches_numpy_golden():
    d = np.load(os.path.join(GOLD, "fft_vectors.npz"))
    keys = sorted(k[3:] for k in d.files if k.startswith("in_"))
    assert len(keys) >= 10
    for key in keys:
        a = d["in_" + key]
        f = O.fft_nd(a, -1)
        b = O.fft_nd(a, +1)
        scale = np.abs(d["fwd_" + key]).max()
        assert np.abs(f - d["fwd_" + key]).max() <= 1e-13 * scale, key
        assert np.abs(b - d["inv_" + key]).max() <= 1e-13 * np.abs(d["inv_" + key]).max(), key


@pytest.mark.parametrize("n", [2, 3, 5, 7, 12, 16, 17, 30, 34, 49, 101])
def test_fft_brute_force_dft(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert np.allclose(O.fft_nd(a, -1), W @ a, rtol=0, atol=1e-12 * np.abs(a).sum())


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("n", [12, 30, 150, 96, 300, 64, 101, 34])
def test_fft_other_radix_orders_are_the_same_dft(n, mode):
    """The other radix orders (used to measure the rounding envelope) are the
    same transform: brute-force DFT, within FP64 rounding; they round
    differently where the factorization changes."""
    rng = np.random.default_rng(n + 1)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    O.set_fft_order(mode)
    try:
        f = O.fft_nd(a, -1)
    finally:
        O.set_fft_order(0)
    assert np.allclose(f, W @ a, rtol=0, atol=1e-12 * np.abs(a).sum())
    changes = (mode == 1 and n in (30, 150, 300, 101, 34)) or (mode == 2 and n in (12, 96, 300, 64))
    if changes:
        assert not np.array_equal(f, O.fft_nd(a, -1))


def test_reversed_deposit_order_is_the_same_binning():
    """The reversed deposit order (a rounding-envelope probe) sums the same
    terms: equal to the default order within FP64 rounding, same support."""
    rng = np.random.default_rng(11)
    g = O.make_grid([-3.0, -3.0], [3.0, 3.0], [33, 29])
    x = rng.uniform(-2.9, 2.9, size=(4000, 2))
    idx = rng.integers(0, 4000, size=(3, 4000)).astype(np.int32)
    _, s0, q0 = O.dirac_sequences(g, x, idx)
    O.lib().og_set_deposit_order(1)
    try:
        _, s1, q1 = O.dirac_sequences(g, x, idx)
    finally:
        O.lib().og_set_deposit_order(0)
    assert np.allclose(s1, s0, rtol=1e-13, atol=0) and np.allclose(q1, q0, rtol=1e-13, atol=0)
    assert np.array_equal(s1 > 0, s0 > 0)
    assert not np.array_equal(s1, s0)  # a different rounding somewhere


def test_fixed_point_deposits_round_the_same_sum():
    """Deposit mode 2 (the fixed-point deterministic mode) rounds each deposit
    to 2^-fb of a whole draw (fb = 61 - bits of n) and sums exactly: each
    node is within (deposits at the node) half-quanta of the FP64 sum, the
    support is the same, and any order of the draws gives the same bits."""
    rng = np.random.default_rng(12)
    g = O.make_grid([-3.0, -3.0, -2.0], [3.0, 3.0, 2.0], [17, 13, 11])
    n = 3000
    x = rng.uniform(-2.9, 2.9, size=(n, 3)) * np.array([1.0, 1.0, 0.6])
    idx = rng.integers(0, n, size=(3, n)).astype(np.int32)
    _, s0, q0 = O.dirac_sequences(g, x, idx)
    fb = O.lib().og_fixed_frac_bits(n)
    assert fb == 61 - int(n).bit_length()
    O.set_deposit_order(2)
    try:
        _, s2, q2 = O.dirac_sequences(g, x, idx)
        _, s2p, q2p = O.dirac_sequences(g, x, idx[:, ::-1].copy())
    finally:
        O.set_deposit_order(0)
    hp = np.prod([g.h[i] for i in range(3)])
    qM = 2.0 ** -fb / (n * hp)
    qV = 2.0 ** -fb / (n * hp) ** 2
    assert np.abs(s2 - s0).max() <= 0.5 * qM * n * 8 and np.abs(q2 - q0).max() <= 0.5 * qV * n * 8
    assert np.array_equal(s2 > 0, s0 > 0)
    assert np.array_equal(s2, s2p) and np.array_equal(q2, q2p)
    # a single deposit is exactly llrint(term * scale) / scale
    one = np.array([[0.123, -0.456, 0.0789]])
    _, s1, _ = O.dirac_sequences(g, one, np.zeros((1, 1), dtype=np.int32))
    O.set_deposit_order(2)
    try:
        _, s1f, _ = O.dirac_sequences(g, one, np.zeros((1, 1), dtype=np.int32))
    finally:
        O.set_deposit_order(0)
    scale = 2.0 ** O.lib().og_fixed_frac_bits(1) * 1 * hp
    nz = s1 > 0
    assert np.array_equal(s1f[nz], np.rint(s1[nz] * scale) * (1.0 / scale))


def test_real_output_inverse_is_the_same_estimate():
    """The real-output probe (og_set_real_output) drops only the imaginary
    rounding noise of the inverse transforms: the log-VMR tables agree to
    rounding where the means are significant (rungs past the first three,
    where empty cells of a resample still hold nothing but that noise), and
    the estimate moves only by the few rungs those first rungs decide."""
    x = W.gaussian_1d(10000, 3)
    g = O.make_grid([-5.0], [5.0], [256])
    idx = O.bootstrap_indices(4, x.shape[0], 20)
    r0 = O.gradepro(g, x, idx_bs=idx, n_bootstraps=20, n_steps=60, tables=True)
    O.lib().og_set_real_output(1)
    try:
        r1 = O.gradepro(g, x, idx_bs=idx, n_bootstraps=20, n_steps=60, tables=True)
    finally:
        O.lib().og_set_real_output(0)
    sig = r0.means >= 1e-2 * r0.means.max(axis=1, keepdims=True)
    sig[:3] = False
    assert np.nanmax(np.abs(r1.vmr - r0.vmr)[sig]) < 1e-8
    assert np.abs(r1.density - r0.density).sum() / r0.density.sum() < 1e-4
    assert np.mean(r1.chosen == r0.chosen) >= 0.98


def test_fft_round_trip_and_parseval():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((6, 10, 9)) + 1j * rng.standard_normal((6, 10, 9))
    f = O.fft_nd(a, -1)
    back = O.fft_nd(f, +1) / a.size
    assert np.abs(back - a).max() < 1e-13
    assert math.isclose((np.abs(f) ** 2).sum(), a.size * (np.abs(a) ** 2).sum(), rel_tol=1e-12)


# ----------------------------------------------------------------- binning
def _reference_kat_sequences(n_dims):
    """directspace_tests.jl:1-41 restated: samples 0.05 and 0.15 on every
    axis, grid -5:0.1:5, indices [1 2; 2 1]."""
    n = 101
    spacing_sq = (0.1 ** n_dims) ** 2
    dens = np.zeros((2,) + (n,) * n_dims)
    dens2 = np.zeros_like(dens)
    import itertools
    for il in (50, 51):  # 0-based low index (Julia 51 / 52)
        for corner in itertools.product((il, il + 1), repeat=n_dims):
            term = (0.05 ** n_dims) / (2 * spacing_sq)
            for b in (0, 1):
                dens[(b,) + tuple(reversed(corner))] += term
                dens2[(b,) + tuple(reversed(corner))] += term ** 2
    return dens.reshape(2, -1), dens2.reshape(2, -1)


@pytest.mark.parametrize("n_dims", [1, 2, 3])
def test_binning_reference_kat(n_dims):
    g = O.make_grid_step([-5.0] * n_dims, [5.0] * n_dims, [0.1] * n_dims, [101] * n_dims)
    samples = np.array([[0.05] * n_dims, [0.15] * n_dims])
    idx = np.array([[0, 1], [1, 0]], dtype=np.int32)  # idx[b][s]
    rc, sums, sqs = O.dirac_sequences(g, samples, idx)
    assert rc == 0
    exp, exp2 = _reference_kat_sequences(n_dims)
    assert np.allclose(sums, exp, rtol=1e-12, atol=0)
    assert np.allclose(sqs, exp2, rtol=1e-12, atol=0)


def _tent_deposit(g, samples, idx_col):
    """Independent formulation of linear binning: each node j receives
    prod_i (h_i - |x_i - node_i|)_+ / (n (prod h)^2)."""
    shape = O.grid_shape(g)
    N = g.ndims
    hs = np.array([g.h[i] for i in range(N)])
    los = np.array([g.lo[i] for i in range(N)])
    nodes = [los[i] + hs[i] * np.arange(shape[i]) for i in range(N)]
    n = len(idx_col)
    out = np.zeros(tuple(reversed(shape)))
    out2 = np.zeros_like(out)
    for s in idx_col:
        x = samples[s]
        ws = [np.clip(hs[i] - np.abs(x[i] - nodes[i]), 0, None) for i in range(N)]
        w = ws[0]
        for i in range(1, N):
            w = np.multiply.outer(ws[i], w)
        term = w / (n * np.prod(hs) ** 2)
        out += term
        out2 += term ** 2
    return out.ravel(), out2.ravel()


@pytest.mark.parametrize("n_dims", [1, 2, 3])
def test_binning_asymmetric_brute_force(n_dims):
    rng = np.random.default_rng(40 + n_dims)
    shape = [13, 9, 7][:n_dims]
    lo = [-1.5, -0.7, -2.0][:n_dims]
    hi = [2.5, 1.3, 1.0][:n_dims]
    g = O.make_grid(lo, hi, shape)
    n = 25
    samples = np.column_stack([rng.uniform(lo[i] + 1e-3, hi[i] - (hi[i] - lo[i]) / (shape[i] - 1) - 1e-3, n)
                               for i in range(n_dims)])
    idx = rng.integers(0, n, size=(3, n)).astype(np.int32)
    rc, sums, sqs = O.dirac_sequences(g, samples, idx)
    assert rc == 0
    for b in range(3):
        e1, e2 = _tent_deposit(g, samples, idx[b])
        assert np.allclose(sums[b], e1, rtol=1e-12, atol=1e-14 * e1.max())
        assert np.allclose(sqs[b], e2, rtol=1e-11, atol=1e-14 * e2.max())


@pytest.mark.parametrize("n_dims", [1, 2, 3])
def test_binning_mass_and_first_moment(n_dims):
    """Linear binning keeps total mass 1/prod(h) and the first moment along
    every axis (catches swapped axes in the weights)."""
    rng = np.random.default_rng(5 + n_dims)
    shape = [40, 30, 20][:n_dims]
    lo = [-3.0, -2.0, -1.0][:n_dims]
    hi = [3.0, 4.0, 1.5][:n_dims]
    g = O.make_grid(lo, hi, shape)
    n = 500
    samples = np.column_stack([rng.uniform(lo[i] + 0.01, hi[i] - 0.2, n) for i in range(n_dims)])
    idx = rng.integers(0, n, size=(2, n)).astype(np.int32)
    rc, sums, _ = O.dirac_sequences(g, samples, idx, with_squares=False)
    hp = np.prod([g.h[i] for i in range(n_dims)])
    for b in range(2):
        f = sums[b].reshape(tuple(reversed(shape)))
        assert math.isclose(f.sum(), 1.0 / hp, rel_tol=1e-12)
        for ax in range(n_dims):
            coords = lo[ax] + g.h[ax] * np.arange(shape[ax])
            axis_np = n_dims - 1 - ax
            bshape = [1] * n_dims
            bshape[axis_np] = shape[ax]
            m1 = (f * coords.reshape(bshape)).sum()
            assert math.isclose(m1, samples[idx[b], ax].mean() / hp, rel_tol=1e-10, abs_tol=1e-10 / hp)


def test_binning_out_of_grid_rejected():
    g = O.make_grid([-1.0], [1.0], [21])
    for bad in (1.0, -1.0, 1.5, np.nextafter(1.0, 0.0)):
        rc, _, _ = O.dirac_sequences(g, np.array([[0.0], [bad]]), np.array([[0, 1]], dtype=np.int32))
        assert rc == 2, bad


# --------------------------------------------------------------------- VMR
@pytest.mark.parametrize("n_dims", [1, 2, 3])
def test_vmr_closed_form(n_dims):
    """directspace_tests.jl:47-61,210-230."""
    d = np.load(os.path.join(GOLD, "vmr_closed_form.npz"))
    m, v, exp = d[f"mean_{n_dims}"], d[f"var_{n_dims}"], d[f"expected_{n_dims}"]
    B = m.shape[0]
    G = int(np.prod(m.shape[1:]))
    out = np.zeros(G)
    sk = np.ascontiguousarray(m.reshape(B, G)).view(np.float64).ravel().copy()
    s2k = np.ascontiguousarray(v.reshape(B, G)).view(np.float64).ravel().copy()
    O.lib().og_scaled_vmr(n_dims, G, B, sk, s2k, np.full(n_dims, 1.0), np.full(n_dims, 0.0), 2, out)
    e = exp.reshape(G)
    fin = np.isfinite(e)
    assert np.array_equal(fin, np.isfinite(out))
    assert np.allclose(out[fin], e[fin], rtol=1e-10, atol=1e-12)


def test_full_means_abs():
    """directspace_tests.jl:63-67,232-245."""
    rng = np.random.default_rng(9)
    z = rng.random(500) + 1j * rng.random(500)
    out = np.zeros(500)
    O.lib().og_full_means(500, z.view(np.float64).copy(), out)
    assert np.allclose(out, np.abs(z), rtol=1e-15)


# ------------------------------------------------------------- convergence
def _snapshot(stage, G, steps_low):
    """directspace_tests.jl:69-128 (CPU snapshots)."""
    if stage == "low_density_recon":
        return (np.full(G, 1.0e6), np.full(G, 0.8), np.full(G, 1.0e6), np.zeros(G),
                np.full(G, steps_low, np.uint16), np.zeros(G, np.uint16), np.zeros(G, np.uint8))
    if stage == "low_density_stop":
        return (np.full(G, 1.0), np.full(G, 0.99), np.full(G, 1.0), np.full(G, np.nan),
                np.full(G, steps_low + 1, np.uint16), np.zeros(G, np.uint16), np.ones(G, np.uint8))
    if stage == "high_density_update":
        return (np.full(G, 1.00005), np.full(G, 1.0), np.full(G, 1.00005), np.zeros(G),
                np.zeros(G, np.uint16), np.zeros(G, np.uint16), np.zeros(G, np.uint8))
    if stage == "high_density_noupdate":
        return (np.full(G, 1.01), np.full(G, 1.0), np.full(G, 1.01), np.full(G, -5.0),
                np.zeros(G, np.uint16), np.zeros(G, np.uint16), np.zeros(G, np.uint8))
    raise ValueError(stage)


@pytest.mark.parametrize("n_dims", [1, 2, 3])
def test_convergence_snapshots(n_dims):
    """directspace_tests.jl:247-368: dlogt 0.1, eps 2, steps_low 3, steps_over 30."""
    G = 101 ** n_dims if n_dims < 3 else 21 ** 3
    dlogt, eps, sl, so = 0.1, 2.0, 3, 30
    L = O.lib()

    def run(stage):
        cur, p1, p2, mins, cl, co, lf = _snapshot(stage, G, sl)
        dens = np.full(G, np.nan)
        means = np.ones(G)
        L.og_identify_convergence(G, dens, means, cur, p1, p2, dlogt, eps, sl, so, mins, cl, co, lf,
                                  None, 0, 0)
        return dens, mins, cl, co, lf

    dens, mins, cl, co, lf = run("low_density_recon")
    assert np.all(np.isnan(dens)) and np.all(np.isnan(mins))
    assert np.all(cl == sl + 1) and np.all(co == 0) and np.all(lf == 1)

    dens, mins, cl, co, lf = run("low_density_stop")
    assert np.all(dens == 1.0) and not np.any(mins == 0)
    assert np.all(cl == sl + 1) and np.all(co == 0) and np.all(lf == 1)

    dens, mins, cl, co, lf = run("high_density_update")
    assert np.all(dens == 1.0) and np.all(mins <= 0)
    assert np.all(cl == 0) and np.all(co == 0) and np.all(lf == 0)

    dens, mins, cl, co, lf = run("high_density_noupdate")
    assert np.all(np.isnan(dens)) and np.all(mins == -5.0)
    assert np.all(cl == 0) and np.all(co == 1) and np.all(lf == 0)


def test_history_rotation():
    """gradepro_tests.jl:127-164: f_prev2 <- f_prev1 <- vmr after a step."""
    G = 50
    vmr = np.full(G, 0.25)
    p1, p2 = np.ones(G), np.full(G, 2.0)
    mins, cl, co, lf = np.full(G, np.nan), np.zeros(G, np.uint16), np.zeros(G, np.uint16), np.zeros(G, np.uint8)
    O.lib().og_identify_convergence(G, np.full(G, np.nan), np.ones(G), vmr, p1, p2, 1e-4, 2.0, 3, 30,
                                    mins, cl, co, lf, None, 0, 1)
    assert np.all(p1 == 0.25) and np.all(p2 == 1.0)


# ------------------------------------------------------------ ladder/grid
def test_ladder_time_step_kat():
    """gradepro_tests.jl:166-177: time_final 2.0, step 0.2."""
    for n_dims in (1, 2, 3):
        bw = np.full(n_dims, 2.0)
        S = O.lib().og_ladder(n_dims, bw, 0, 0.2, None, None)
        times = np.zeros(S * n_dims)
        dt = np.zeros(n_dims)
        O.lib().og_ladder(n_dims, bw, 0, 0.2, dt.ctypes.data, times.ctypes.data)
        times = times.reshape(S, n_dims)
        assert np.all(times[0] == 0.0) and np.all(times[-1] == 2.0) and np.all(dt == 0.2)
    assert O.lib().og_ladder(1, np.array([1.0]), 0, 1.5, None, None) == -1


def test_ladder_defaults_and_duration_steps():
    """GradeProEstimator.jl:754-766,954-966: (steps_low, steps_over) is
    (0, 25) in 1-D and (15, 25) in 2-D/3-D with 250 steps."""
    for n_dims, alpha_s, expect in ((1, 0.0, (0, 25)), (2, 0.06, (15, 25)), (3, 0.06, (15, 25))):
        bw = np.linspace(0.11, 0.37, n_dims)
        S = O.lib().og_ladder(n_dims, bw, 250, np.nan, None, None)
        assert S == 251
        times = np.zeros(S * n_dims)
        dt = np.zeros(n_dims)
        O.lib().og_ladder(n_dims, bw, 250, np.nan, dt.ctypes.data, times.ctypes.data)
        import ctypes as C
        sl, so = C.c_int(), C.c_int()
        O.lib().og_duration_steps(n_dims, times[-n_dims:].copy(), dt, alpha_s, 0.1, C.byref(sl), C.byref(so))
        assert (sl.value, so.value) == expect


def test_dlogt_first_rung_is_zero():
    t0 = np.array([0.05, 0.02])
    assert O.lib().og_dlogt(2, t0, np.zeros(2), None) == 0.0
    t1, t2 = np.array([0.1, 0.2]), np.array([0.2, 0.4])
    expect = math.log((0.0025 + 0.04) * (0.0004 + 0.16) / ((0.0025 + 0.01) * (0.0004 + 0.04)))
    assert math.isclose(O.lib().og_dlogt(2, t0, t2, t1.ctypes.data), expect, rel_tol=1e-14)


def test_fftgrid_matches_numpy():
    """grid_tests.jl:22-27: fftgrid = 2pi fftfreq(n, 1/h)."""
    for n, h in ((21, 0.1), (512, 10 / 511), (300, 12 / 299), (7, 0.5)):
        assert np.allclose(O.fftfreq(n, h), 2 * np.pi * np.fft.fftfreq(n, h), rtol=1e-15, atol=1e-15)


def test_silverman_closed_form():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1000, 2)) * np.array([1.0, 3.0])
    bw = O.silverman(x)
    pref = (4.0 / (4 * 1000)) ** (1 / 6)
    assert np.allclose(bw, pref * x.std(axis=0, ddof=1), rtol=1e-13)


# ------------------------------------------------------------- propagation
def _direct_kde(samples, nodes, sigma):
    """Direct Gaussian KDE on the grid and its squared-kernel companion
    (fourierspace_tests.jl:1-25)."""
    n, N = samples.shape
    dens = 0.0
    dens2 = 0.0
    for s in range(n):
        k = 1.0
        for i in range(N):
            k = np.multiply.outer(np.exp(-0.5 * ((nodes[i] - samples[s, i]) / sigma) ** 2) /
                                  (np.sqrt(2 * np.pi) * sigma), k) if i > 0 else \
                np.exp(-0.5 * ((nodes[i] - samples[s, i]) / sigma) ** 2) / (np.sqrt(2 * np.pi) * sigma)
        k = k / n
        dens = dens + k
        dens2 = dens2 + k ** 2
    return np.asarray(dens), np.asarray(dens2)


@pytest.mark.parametrize("n_dims", [1, 2])
def test_propagation_gaussian_semigroup(n_dims):
    """fourierspace_tests.jl:27-96: a KDE at sigma=0.2 propagated by t=0.2
    equals the KDE at sqrt(2)*0.2; the squared field within rtol 1e-6."""
    rng = np.random.default_rng(100 + n_dims)
    samples = rng.standard_normal((100, n_dims))
    n = 101
    g = O.make_grid_step([-5.0] * n_dims, [5.0] * n_dims, [0.1] * n_dims, [n] * n_dims)
    nodes = [np.linspace(-5.0, 5.0, n)] * n_dims
    d0, d20 = _direct_kde(samples, nodes, 0.2)
    d1, d21 = _direct_kde(samples, nodes, np.sqrt(2 * 0.2 ** 2))
    f0 = O.fft_nd(d0.astype(complex), -1)
    f20 = O.fft_nd(d20.astype(complex), -1)
    kg = np.concatenate([O.fftfreq(n, 0.1)] * n_dims)
    G = n ** n_dims
    mt = np.zeros(2 * G)
    vt = np.zeros(2 * G)
    t = np.full(n_dims, 0.2)
    import ctypes
    m0 = f0.ravel().view(np.float64).copy()
    v0 = f20.ravel().view(np.float64).copy()
    O.lib().og_propagate(ctypes.byref(g), kg, t, t, m0, v0.ctypes.data, mt, vt.ctypes.data)
    m = mt.view(np.complex128).reshape(d0.shape)
    v = vt.view(np.complex128).reshape(d0.shape)
    f1 = O.fft_nd(d1.astype(complex), -1)
    f21 = O.fft_nd(d21.astype(complex), -1)
    assert np.allclose(m, f1, rtol=1e-6, atol=1e-8 * np.abs(f1).max())
    assert np.allclose(v, f21, rtol=1e-6, atol=1e-6 * np.abs(f21).max())
    back = np.abs(O.fft_nd(m, +1) / G)
    assert np.allclose(back, d1, rtol=1e-6, atol=1e-8 * d1.max())


# -------------------------------------------------------------- end to end
def test_end_to_end_mise_1d():
    """gradepro_tests.jl:485-511 (N=1): 10,000 N(0,1) samples, grid
    range(-6, 6, length=300), MISE < 5e-5 (seeded here)."""
    from scipy.stats import norm
    rng = np.random.default_rng(1)
    x = rng.standard_normal((10000, 1))
    g = O.make_grid([-6.0], [6.0], [300])
    r = O.gradepro(g, x, seed=5, threads=4)
    assert r.rc == 0 and r.info.n_rungs == 251
    xs = np.linspace(-6, 6, 300)
    mise = ((r.density - norm.pdf(xs)) ** 2).sum() * (xs[1] - xs[0]) / 300
    assert mise < 5e-5
    assert np.all(np.isfinite(r.density))


def test_serial_equals_threaded():
    rng = np.random.default_rng(8)
    x = rng.standard_normal((3000, 2)) * 0.8
    g = O.make_grid([-5.0, -5.0], [5.0, 5.0], [24, 20])
    a = O.gradepro(g, x, seed=3, n_bootstraps=16, n_steps=40, threads=1)
    b = O.gradepro(g, x, seed=3, n_bootstraps=16, n_steps=40, threads=4)
    assert np.array_equal(a.density, b.density) and np.array_equal(a.chosen, b.chosen)


def test_chosen_rung_consistent_with_means_table():
    """The chosen rung reproduces the density from the per-rung means."""
    rng = np.random.default_rng(12)
    x = rng.standard_normal((2000, 1))
    g = O.make_grid([-5.0], [5.0], [128])
    r = O.gradepro(g, x, seed=1, n_bootstraps=20, n_steps=60, tables=True)
    S = r.info.n_rungs
    sel = np.where(r.chosen == S, S - 1, r.chosen)
    assert np.array_equal(r.density, r.means[sel, np.arange(128)])


# ------------------------------------------------------- :rot estimator
def _mise(f1, f2, dx):
    """calculate_mise (test/test_utils.jl:24-28)."""
    return ((f1 - f2) ** 2).sum() * dx / f1.size


@pytest.mark.parametrize("n_dims", [1, 2])
@pytest.mark.parametrize("rule", [0, 1])
def test_rot_mise_reference_test(n_dims, rule):
    """rulesofthumb_tests.jl:96-126: 1000 N(0,I) samples, grid
    range(-6, 6, length=300) per axis, MISE < 1e-3 against the true pdf
    (seeded here; both rules)."""
    from scipy.stats import norm
    rng = np.random.default_rng(60 + n_dims)
    x = rng.standard_normal((1000, n_dims))
    g = O.make_grid([-6.0] * n_dims, [6.0] * n_dims, [300] * n_dims)
    rc, dens, t = O.rot(g, x, rule)
    assert rc == 0
    xs = np.linspace(-6, 6, 300)
    truth = norm.pdf(xs)
    if n_dims == 2:
        truth = np.outer(truth, truth).ravel(order="F")
    dx = (xs[1] - xs[0]) ** n_dims
    assert _mise(dens, truth, dx) < 1e-3
    assert np.all(np.isfinite(dens))


@pytest.mark.parametrize("n_dims", [1, 2, 3])
def test_rot_propagation_time(n_dims):
    """propagation_time (RoTEstimator.jl:22-28, 42-47) minus the initial
    bandwidth h/2 (RoTEstimator.jl:67-70), against numpy's ddof=1 standard
    deviation: Scott n^(-1/(N+4)), Silverman (n (N+2)/4)^(-1/(N+4))."""
    rng = np.random.default_rng(70 + n_dims)
    n = 777
    x = rng.standard_normal((n, n_dims)) * np.array([1.0, 0.5, 2.0][:n_dims]) + 0.3
    g = O.make_grid([-9.0] * n_dims, [9.0] * n_dims, [32] * n_dims)
    sd = x.std(axis=0, ddof=1)
    h = np.array([g.h[i] for i in range(n_dims)])
    for rule, base in ((0, float(n)), (1, n * (n_dims + 2) / 4.0)):
        rc, _, t = O.rot(g, x, rule)
        assert rc == 0
        assert np.allclose(t, base ** (-1.0 / (n_dims + 4)) * sd - h / 2, rtol=1e-12, atol=0)


@pytest.mark.parametrize("n_dims", [1, 2, 3])
def test_rot_matches_numpy_pipeline(n_dims):
    """The oracle's :rot after the Dirac sequence (pinned by the binning KAT)
    against numpy's FFT: |ifftn(fftn(f) * exp(-sum (k_i t_i)^2 / 2))| with
    k = 2 pi fftfreq(n, h) (FourierSpace.jl:237-262; RoTEstimator.jl:
    237-265), odd and even axis lengths."""
    rng = np.random.default_rng(80 + n_dims)
    x = np.clip(rng.standard_normal((500, n_dims)), -3.0, 3.0)
    shape = [21, 16, 12][:n_dims]
    g = O.make_grid([-5.0] * n_dims, [5.0] * n_dims, shape)
    rc, dens, t = O.rot(g, x, 1)
    assert rc == 0
    _, f, _ = O.dirac_sequences(g, x, np.arange(500, dtype=np.int32)[None, :])
    F = np.fft.fftn(f[0].reshape(shape, order="F"))
    expo = np.zeros(shape)
    for i in range(n_dims):
        k = 2 * np.pi * np.fft.fftfreq(shape[i], g.h[i])
        sh = [1] * n_dims
        sh[i] = shape[i]
        expo = expo + (k * t[i]).reshape(sh) ** 2
    ref = np.abs(np.fft.ifftn(F * np.exp(-0.5 * expo))).ravel(order="F")
    assert np.allclose(dens, ref, rtol=1e-11, atol=1e-13 * ref.max())


def test_rot_out_of_grid():
    g = O.make_grid([-1.0], [1.0], [16])
    rc, _, _ = O.rot(g, np.array([[0.0], [2.0]]), 0)
    assert rc == 2
