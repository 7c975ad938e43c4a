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
