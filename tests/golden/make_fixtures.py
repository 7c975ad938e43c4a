"""Writes the golden vectors under tests/golden/ that pin the oracle.

The reference ships no golden files (every expected value in its test suite is
computed in-line), and Julia/FFTW are not available here.  These fixtures come
from independent implementations available in this container:

* fft_vectors.npz   numpy.fft (pocketfft) forward/inverse C2C transforms of
                    seeded complex arrays at the sizes the reference's tests
                    and the benchmark configurations use (1-D 8..512 incl.
                    21/101/300, 2-D, 3-D).
* vmr_closed_form.npz  the closed form of directspace_tests.jl:47-61 computed
                    with numpy's two-pass variance (ddof=1) on seeded complex
                    arrays of shape 50^N x 5, N = 1..3 (t = 1, t0 = 0, n = 2).
* philox_kat.json   Philox4x32-10 known-answer vectors published with
                    Random123 (kat_vectors), used by the bootstrap RNG.

Run:  python tests/golden/make_fixtures.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_vectors():
    rng = np.random.default_rng(20240611)
    out = {}
    shapes = [(8,), (16,), (21,), (64,), (101,), (128,), (300,), (512,),
              (8, 12), (16, 16), (21, 10), (4, 6, 5), (8, 8, 8)]
    for k, shp in enumerate(shapes):
        a = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        key = "x".join(str(s) for s in shp)
        out[f"in_{key}"] = a
        out[f"fwd_{key}"] = np.fft.fftn(a)
        out[f"inv_{key}"] = np.fft.ifftn(a) * a.size  # unnormalised inverse
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


def vmr_closed_form():
    rng = np.random.default_rng(7)
    out = {}
    for n_dims in (1, 2, 3):
        side = 50 if n_dims < 3 else 16
        shp = (5,) + (side,) * n_dims  # numpy order: bootstrap axis first
        m = rng.random(shp) + 1j * rng.random(shp)
        v = rng.random(shp) + 1j * rng.random(shp)
        n_samples = 2
        means = np.abs(m) / n_samples
        vars_ = np.abs(v) / n_samples - means ** 2
        vmr = vars_ / means
        scaling = np.prod(np.full(n_dims, 1.0) ** 2 + np.full(n_dims, 0.0) ** 2) ** 1.5 * n_samples ** 3
        vv = scaling * np.var(vmr, axis=0, ddof=1)
        with np.errstate(divide="ignore", invalid="ignore"):
            expected = np.where(np.isfinite(vv), np.log(vv), np.nan)
        out[f"mean_{n_dims}"] = m
        out[f"var_{n_dims}"] = v
        out[f"expected_{n_dims}"] = expected
    np.savez_compressed(os.path.join(HERE, "vmr_closed_form.npz"), **out)


def philox_kat():
    # Random123 kat_vectors, philox4x32 with 10 rounds: (ctr, key, expected)
    kat = [
        [[0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]],
        [[0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]],
        [[0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]],
    ]
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)


if __name__ == "__main__":
    fft_vectors()
    vmr_closed_form()
    philox_kat()
    print("fixtures written to", HERE)
