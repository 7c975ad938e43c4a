"""CPU-side tests of the boundary: the C-ABI library loads and exports every
entry point its header declares, and the Python mirror of the reference API
behaves like ParallelKDE.jl's host logic (grid construction, validation,
error mapping).  No GPU compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "parallelkde_hip.h")


def _declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pkde_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    from parallelkde_rocm import _lib
    L = _lib.load()
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(_lib.EXPORTED_SYMBOLS)
    text = open(HEADER).read()
    abi = int(re.search(r"#define PKDE_ABI_VERSION (\d+)", text).group(1))
    dims = int(re.search(r"#define PKDE_MAX_DIMS (\d+)", text).group(1))
    assert L.pkde_version() == abi == _lib.ABI_VERSION
    assert dims == _lib.MAX_DIMS  # the ctypes structs' per-axis arrays


def test_ctypes_structs_match_header(tmp_path):
    """The Python mirror of the ABI structs has the C layout: sizes and field
    offsets printed by a C program compiled against the header."""
    import shutil
    import subprocess
    from parallelkde_rocm import _lib
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    checks = [("pkde_gradepro_params", _lib.Params, ["bw_final", "alpha_s", "rng_seed", "precision"]),
              ("pkde_stats", _lib.Stats, ["eps", "bw_final", "ms_total", "algorithmic_bytes"]),
              ("pkde_request", _lib.Request, ["n_per_axis", "h", "samples", "params", "chosen_out"]),
              ("pkde_server_cfg", _lib.ServerCfg, ["memory_budget"])]
    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, _, fields in checks:
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        lines += [f'  printf("%zu\\n", offsetof({cname}, {f}));' for f in fields]
    lines += ["  return 0;", "}"]
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "abi"
    subprocess.run([cc, "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    want = []
    for _, cls, fields in checks:
        want.append(C.sizeof(cls))
        want += [getattr(cls, f).offset for f in fields]
    assert got == want


def test_params_default_matches_reference_defaults():
    from parallelkde_rocm import _lib
    p = _lib.Params()
    _lib.load().pkde_gradepro_params_default(C.byref(p))
    assert p.n_bootstraps == 100 and p.n_steps == 250  # GradeProEstimator.jl:705, :902
    assert np.isnan(p.bw_step) and all(np.isnan(p.bw_final[i]) for i in range(3))
    assert np.isnan(p.alpha_s) and np.isnan(p.alpha_os) and np.isnan(p.eps)
    assert p.deterministic_binning == 1 and p.precision == 1


def test_no_device_fails_loudly():
    import parallelkde_rocm as pk
    from parallelkde_rocm import _lib
    h = C.c_void_p()
    rc = _lib.load().pkde_context_create(0, C.byref(h))
    if rc == 0:
        _lib.load().pkde_context_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == _lib.NO_DEVICE
    est = pk.initialize_estimation(np.random.default_rng(0).standard_normal((50, 1)),
                                   grid=pk.initialize_grid([(-5.0, 5.0, 64)]))
    with pytest.raises(pk.DeviceUnavailableError):
        pk.estimate_density(est, "gradepro")


def test_grid_geometry_kat():
    """grid_tests.jl:1-36: sizes, spacings, bounds, fftgrid = 2pi fftfreq."""
    import parallelkde_rocm as pk
    for n_dims in (1, 2, 3):
        g = pk.initialize_grid([(-1.0, 1.0, 21)] * n_dims)
        assert g.shape == (21,) * n_dims and g.ndims == n_dims
        assert np.allclose(pk.spacings(g), 0.1, rtol=1e-15)
        assert np.array_equal(pk.bounds(g), np.array([[-1.0] * n_dims, [1.0] * n_dims]))
        assert np.array_equal(pk.low_bounds(g), [-1.0] * n_dims)
        assert np.array_equal(pk.high_bounds(g), [1.0] * n_dims)
        assert np.allclose(pk.initial_bandwidth(g), 0.05)
        for k in pk.fftgrid(g):
            assert np.allclose(k, 2 * np.pi * np.fft.fftfreq(21, 0.1), rtol=1e-14, atol=1e-14)
        coords = g.get_coordinates()
        assert coords.shape == (n_dims,) + (21,) * n_dims


def test_find_grid_padding():
    """grid_tests.jl:29-33: data at +-1 gives bounds +-1.2 (10 % padding)."""
    import parallelkde_rocm as pk
    for n_dims in (1, 2, 3):
        data = np.array([[-1.0] * n_dims, [1.0] * n_dims])
        g = pk.find_grid(data)
        assert np.allclose(pk.bounds(g), np.array([[-1.2] * n_dims, [1.2] * n_dims]), rtol=1e-15)
        assert g.shape == (pk.FIND_GRID_DEFAULT_POINTS,) * n_dims


def test_initialize_estimation_contract():
    """ParallelKDE.jl:102-158 and KDEs.jl:62-111."""
    import parallelkde_rocm as pk
    x = np.random.default_rng(1).standard_normal((100, 2))
    est = pk.initialize_estimation(x, grid=True, grid_ranges=[(-5.0, 5.0, 64), (-4.0, 4.0, 32)])
    assert pk.has_grid(est) and pk.get_grid(est).shape == (64, 32)
    assert est.kde.density.shape == (64, 32) and np.all(np.isnan(est.kde.density))
    assert est.kde.get_nsamples() == 100
    with pytest.raises(pk.ArgumentError):
        pk.initialize_estimation(x, grid=False)  # no grid and no dims
    est2 = pk.initialize_estimation(x, grid=False, dims=(8, 8))
    assert not pk.has_grid(est2)
    with pytest.raises(pk.ArgumentError):
        pk.initialize_estimation(x, grid=True, device=":cuda")  # not a device of this package
    with pytest.raises(pk.ArgumentError):
        pk.initialize_kde(x, (10, 10, 10))  # dimension mismatch


def test_estimate_argument_validation():
    import parallelkde_rocm as pk
    x = np.random.default_rng(2).standard_normal((100, 1))
    est = pk.initialize_estimation(x, grid=pk.initialize_grid([(-5.0, 5.0, 64)]))
    with pytest.raises(KeyError):
        pk.estimate_density(est, "nosuchestimator")
    with pytest.raises(pk.ArgumentError):
        pk.estimate_density(est, "gradepro", method="serial")  # ensure_valid_implementation
    with pytest.raises(TypeError):
        pk.estimate_density(est, "gradepro", bandwidth=0.3)


def test_get_density_normalisation():
    """ParallelKDE.jl:189-202: normalise in place by sum * prod(h)."""
    import parallelkde_rocm as pk
    x = np.random.default_rng(3).standard_normal((10, 1))
    est = pk.initialize_estimation(x, grid=pk.initialize_grid([(-5.0, 5.0, 11)]))
    est.kde.density[:] = 2.0
    d = pk.get_density(est, normalize=True)
    assert np.isclose(d.sum() * 1.0, 1.0)
    est3 = pk.initialize_estimation(x, grid=False, dims=(11,))
    with pytest.raises(pk.ArgumentError):
        pk.get_density(est3, normalize=True)
    est3.kde.density[:] = 1.0
    assert np.isclose(pk.get_density(est3, normalize=True, dx=0.5).sum(), 2.0)


JULIA_GLUE = os.path.join(ROOT, "parallelkde.jl_amd", "julia", "ROCmBackend.jl")


def test_julia_glue_struct_declarations_are_definable():
    """Julia cannot compute with a type variable in a supertype (`... <:
    Abstract{N, T, N + 1}` throws at definition time), so every struct of the
    glue names each supertype parameter directly (GradeProEstimator.jl:673
    carries the bootstrapped rank M as a parameter of its own)."""
    text = open(JULIA_GLUE).read()
    decls = re.findall(r"^(?:mutable\s+)?struct\s+([^\n]+)$", text, flags=re.M)
    assert decls
    for d in decls:
        parts = re.match(r"^(\w+)(?:\{([^}]*)\})?\s+<:\s+(.*)$", d.strip())
        if not parts:
            continue
        own, sup = parts.group(2), parts.group(3)
        params = {p.split("<:")[0].strip() for p in own.split(",")} if own else set()
        m = re.search(r"\{([^}]*)\}\s*$", sup.strip())
        if not m:
            continue
        for arg in (a.strip() for a in m.group(1).split(",")):
            assert re.fullmatch(r"[A-Za-z_][A-Za-z0-9_.]*", arg), (d, arg)
            if len(arg) <= 2 and arg[0].isupper():
                assert arg in params, (d, arg)


def test_julia_glue_falls_back_to_cpu():
    """The :rocm constructors warn and fall back to :cpu without a functional
    device, as ParallelKDE.jl:120-128, Grids.jl:41-44 and KDEs.jl:68-71 do for
    :cuda."""
    text = open(JULIA_GLUE).read()
    for fn in ("Grids.initialize_grid(::IsROCm", "KDEs.initialize_kde(::IsROCm"):
        i = text.index(fn)
        body = text[i:text.index("\nend", i)]
        assert "!functional()" in body and "warn_fallback()" in body and "IsCPU()" in body, fn
    assert "Falling back to ':cpu'" in text


def test_julia_params_mirror_c_struct():
    """PkdeParams in the glue lists the fields of pkde_gradepro_params in the
    header's order with the same widths."""
    text = open(JULIA_GLUE).read()
    jl = re.search(r"struct PkdeParams\n(.*?)\nend", text, flags=re.S).group(1)
    jfields = [ln.split("::")[0].strip() for ln in jl.strip().splitlines()]
    h = open(HEADER).read()
    cs = re.search(r"typedef struct \{(.*?)\} pkde_gradepro_params;", h, flags=re.S).group(1)
    cs = re.sub(r"/\*.*?\*/", "", cs, flags=re.S)
    cfields = [re.sub(r"\[.*\]", "", ln.split()[-1]).rstrip(";") for ln in cs.strip().splitlines()
               if ln.strip()]
    assert jfields == cfields


def test_integration_notes_state_the_header_abi():
    h = open(HEADER).read()
    abi = int(re.search(r"#define PKDE_ABI_VERSION (\d+)", h).group(1))
    notes = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    stated = {int(v) for v in re.findall(r"ABI version[^0-9]{0,40}?(\d+)", notes)}
    assert stated == {abi}, stated


def _build_c_caller(tmp_path):
    """Compiles tests/c_abi_caller.c with the host C compiler against the
    public header and the built library (what a C host does)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    libdir = os.path.join(ROOT, "parallelkde.jl_amd", "lib")
    exe = tmp_path / "c_abi_caller"
    subprocess.run([cc, "-O2", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", str(exe), os.path.join(ROOT, "tests", "c_abi_caller.c"),
                    "-L", libdir, "-lparallelkde_hip", f"-Wl,-rpath,{libdir}", "-lm"], check=True)
    return exe


def test_c_caller_builds_and_reports_no_device(tmp_path):
    """A plain C program compiles against include/parallelkde_hip.h and links
    the library; without a GPU it gets PKDE_NO_DEVICE from
    pkde_context_create and exits 3 (the glue's cue for the :cpu fallback)."""
    import subprocess
    from parallelkde_rocm import _lib
    h = C.c_void_p()
    rc = _lib.load().pkde_context_create(0, C.byref(h))
    if rc == 0:
        _lib.load().pkde_context_destroy(h)
        pytest.skip("a GPU is present (the GPU test runs the program)")
    exe = _build_c_caller(tmp_path)
    r = subprocess.run([str(exe), str(tmp_path / "out.bin")], capture_output=True, text=True)
    assert r.returncode == 3, (r.returncode, r.stdout, r.stderr)
    assert "no usable GPU" in r.stdout
