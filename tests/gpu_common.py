"""Shared helpers of the GPU parity tests: synthetic workloads and the
comparison metrics against the oracle."""
from __future__ import annotations

import numpy as np


def gaussian_1d(n, seed, lo=-5.0, hi=5.0, margin=0.05):
    """N(0,1) draws; draws outside (lo, hi - margin) are redrawn (the last
    cell below hi is excluded by the upper-edge rule)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n)
    bad = (x <= lo) | (x >= hi - margin)
    while bad.any():
        x[bad] = rng.standard_normal(bad.sum())
        bad = (x <= lo) | (x >= hi - margin)
    return x[:, None]


def mixture_2d(n, seed, lo=-6.0, hi=6.0, margin=0.06):
    """Equal-weight 3-component mixture, means (-2,0),(2,0),(0,2.5), sd 0.6."""
    rng = np.random.default_rng(seed)
    means = np.array([[-2.0, 0.0], [2.0, 0.0], [0.0, 2.5]])
    c = rng.integers(0, 3, n)
    x = means[c] + 0.6 * rng.standard_normal((n, 2))
    bad = np.any((x <= lo) | (x >= hi - margin), axis=1)
    while bad.any():
        c = rng.integers(0, 3, bad.sum())
        x[bad] = means[c] + 0.6 * rng.standard_normal((bad.sum(), 2))
        bad = np.any((x <= lo) | (x >= hi - margin), axis=1)
    return x


def correlated_3d(n, seed, lo=-4.0, hi=4.0, margin=0.07, background=0.1):
    """Correlated Gaussian (unit variances, rho 0.5) plus a uniform background
    over the box (background fraction of the n samples)."""
    rng = np.random.default_rng(seed)
    cov = 0.5 * np.eye(3) + 0.5
    Lc = np.linalg.cholesky(cov)
    nb = int(round(background * n))
    ng = n - nb
    g = rng.standard_normal((ng, 3)) @ Lc.T
    bad = np.any((g <= lo) | (g >= hi - margin), axis=1)
    while bad.any():
        g[bad] = rng.standard_normal((bad.sum(), 3)) @ Lc.T
        bad = np.any((g <= lo) | (g >= hi - margin), axis=1)
    u = rng.uniform(lo + 1e-6, hi - margin, (nb, 3))
    x = np.concatenate([g, u])
    return x[rng.permutation(n)]


def gaussian_nd(n, d, seed, lo=-4.0, hi=4.0, margin=0.07, rho=0.3):
    """N(0, S) in d dimensions, S = (1 - rho) I + rho 11^T; draws outside the
    box (lo, hi - margin) on any axis are redrawn."""
    rng = np.random.default_rng(seed)
    Lc = np.linalg.cholesky((1.0 - rho) * np.eye(d) + rho)
    x = rng.standard_normal((n, d)) @ Lc.T
    bad = np.any((x <= lo) | (x >= hi - margin), axis=1)
    while bad.any():
        x[bad] = rng.standard_normal((bad.sum(), d)) @ Lc.T
        bad = np.any((x <= lo) | (x >= hi - margin), axis=1)
    return x


def compare(dens_gpu, chosen_gpu, dens_ref, chosen_ref, S, floor=1e-3):
    """Relative L1 / Linf density error and chosen-rung agreement, overall and
    on points whose oracle density is at least `floor` of the maximum."""
    dg = np.asarray(dens_gpu).ravel(order="F")
    cg = np.asarray(chosen_gpu).ravel(order="F")
    rel_l1 = float(np.abs(dg - dens_ref).sum() / np.abs(dens_ref).sum())
    rel_linf = float(np.abs(dg - dens_ref).max() / np.abs(dens_ref).max())
    sig = dens_ref >= floor * dens_ref.max()
    agree_all = float(np.mean(cg == chosen_ref))
    agree_sig = float(np.mean(cg[sig] == chosen_ref[sig]))
    return dict(rel_l1=rel_l1, rel_linf=rel_linf, agree_all=agree_all, agree_sig=agree_sig,
                n_sig=int(sig.sum()))


def box_draws(rng, n, d, lo, hi, kind, margin):
    """N(0,I) or the mixture 1/2 N(-1.5*1, 0.5^2 I) + 1/2 N(+1.5*1, I) in d
    dimensions; draws outside (lo, hi - margin) on any axis are redrawn."""
    def draw(m):
        if kind == "normal":
            return rng.standard_normal((m, d))
        comp = rng.integers(0, 2, m)[:, None]
        return np.where(comp == 0, -1.5 + 0.5 * rng.standard_normal((m, d)),
                        1.5 + rng.standard_normal((m, d)))
    x = draw(n)
    bad = np.any((x <= lo) | (x >= hi - margin), axis=1)
    while bad.any():
        x[bad] = draw(int(bad.sum()))
        bad = np.any((x <= lo) | (x >= hi - margin), axis=1)
    return x


def batch_mix_items(count=64):
    """The 64-dataset batched workload: dataset j (seed 1000 + j) has a
    dimension uniform in {1, 2, 3}, a sample count log-uniform in [1e3, 2e5],
    N(0,I) or 2-component mixture draws, and a grid of 1024 (1-D), 128^2
    (2-D) or 64^3 (3-D) points over [-6, 6] per axis.  Returns
    [(samples (n, d), [(lo, hi, points)] * d, seed)]."""
    import math
    items = []
    for j in range(count):
        rng = np.random.default_rng(1000 + j)
        d = int(rng.integers(1, 4))
        n = int(10 ** rng.uniform(3, math.log10(2e5)))
        kind = "normal" if rng.random() < 0.5 else "mixture"
        pts = {1: 1024, 2: 128, 3: 64}[d]
        x = box_draws(rng, n, d, -6.0, 6.0, kind, 12.0 / (pts - 1) + 1e-9)
        items.append((x, [(-6.0, 6.0, pts)] * d, 1000 + j))
    return items


def stream_requests(count, seed=7):
    """Requests of the served Poisson stream: request r (seed
    seed * 100000 + r) has a dimension uniform in {1, 2, 3}, a sample count
    log-uniform in [1e3, 1e6] and per-axis points log-uniform in 256-1024
    (1-D), 64-256 (2-D), 32-128 (3-D), rounded to the nearest power of two;
    N(0,I) draws on [-6, 6]^d.  Returns [(samples, [(lo, hi, points)] * d,
    seed)]."""
    import math
    reqs = []
    for r in range(count):
        s = seed * 100000 + r
        rng = np.random.default_rng(s)
        d = int(rng.integers(1, 4))
        n = int(10 ** rng.uniform(3, 6))
        lo_p, hi_p = {1: (256, 1024), 2: (64, 256), 3: (32, 128)}[d]
        pts = int(2 ** round(rng.uniform(math.log2(lo_p), math.log2(hi_p))))
        x = box_draws(rng, n, d, -6.0, 6.0, "normal", 12.0 / (pts - 1) + 1e-9)
        reqs.append((x, [(-6.0, 6.0, pts)] * d, s))
    return reqs


def mass(density, spacing):
    """Integral of a grid density: sum times the cell volume."""
    return float(np.sum(density) * np.prod(spacing))


def check_properties(density, chosen, S, spacing, mass_tol=1e-3):
    """Size-independent properties of a :gradepro density: finite,
    non-negative, unit mass (the binning puts mass 1 on the grid and every
    Gaussian propagation keeps it), chosen rung in [0, S] (S = fallback)."""
    d = np.asarray(density)
    c = np.asarray(chosen)
    assert np.all(np.isfinite(d)), "non-finite density"
    assert np.all(d >= 0), "negative density"
    m = mass(d, spacing)
    assert abs(m - 1.0) <= mass_tol, m
    assert c.min() >= 0 and c.max() <= S, (c.min(), c.max())
    return m
