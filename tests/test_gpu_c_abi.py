"""A plain C caller of the library on the GPU (tests/c_abi_caller.c, built
with the host C compiler against include/parallelkde_hip.h): its own checks
(plain form properties, batched and served results equal to the plain
form's bit for bit, out-of-grid and unsupported-size errors) must pass, and
the 1-D and 2-D densities it computed must match the FP64 oracle run on the
samples it drew, with the same Philox draws (og_bootstrap_indices(seed)).
Bars as in the full-width parity tests: relative L1 <= 1e-9 and every
point whose density is >= 1e-3 of the maximum on the oracle's rung, or,
where the oracle's own result moves further under FP64 rounding alone, 4x
its relative L1 and 3x its significant-point disagreement."""
import os
import subprocess
import sys

import numpy as np
import pytest

import gpu_common as W
import oracle_ctypes as O
from test_gpu_parity import _check_fp64, _rounding_envelope
from test_host import _build_c_caller

pytestmark = pytest.mark.gpu


def _read_cases(path):
    cases = []
    with open(path, "rb") as f:
        while True:
            hdr = f.read(4)
            if not hdr:
                break
            d = int(np.frombuffer(hdr, np.int32)[0])
            n, seed, G = (int(v) for v in np.frombuffer(f.read(24), np.int64))
            x = np.frombuffer(f.read(8 * d * n), np.float64).reshape(n, d)
            dens = np.frombuffer(f.read(8 * G), np.float64)
            chosen = np.frombuffer(f.read(2 * G), np.int16)
            cases.append((d, n, seed, x, dens, chosen))
    return cases


def test_c_caller_matches_oracle(tmp_path):
    exe = _build_c_caller(tmp_path)
    out = tmp_path / "cases.bin"
    print("[c-abi] running the C caller", flush=True)
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    sys.stdout.write(r.stderr)
    assert r.returncode == 0, r.returncode
    cases = _read_cases(out)
    assert [c[0] for c in cases] == [1, 2]
    grids = {1: ([-5.0], [5.0], [512]), 2: ([-6.0] * 2, [6.0] * 2, [64] * 2)}
    for d, n, seed, x, dens, chosen in cases:
        og = O.make_grid(*grids[d])
        ref = O.gradepro(og, np.ascontiguousarray(x), seed=seed, threads=16)
        assert ref.rc == 0
        m = W.compare(dens, chosen, ref.density, ref.chosen, 251)
        print(f"  C caller {d}-D vs oracle: rel L1 {m['rel_l1']:.3e} rel Linf "
              f"{m['rel_linf']:.3e} chosen agree {m['agree_all']:.5f} (sig {m['agree_sig']:.5f} "
              f"of {m['n_sig']})", flush=True)
        idx = O.bootstrap_indices(seed, n, 100)
        _check_fp64(m, lambda: _rounding_envelope(og, np.ascontiguousarray(x), idx, 251,
                                                  deposits=True, n_bootstraps=100, n_steps=250))
