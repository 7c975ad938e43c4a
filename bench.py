"""Benchmark of the :gradepro hot path on one MI355X.

Default workload (the 2-D configuration the headline metric is quoted on):
  100,000 samples from an equal-weight 3-component Gaussian mixture (means
  (-2,0), (2,0), (0,2.5), isotropic sd 0.6, seed 2), 256 x 256 grid over
  [-6,6]^2, B = 100 bootstraps, 251 bandwidth rungs, FP32 spectra and FFTs
  with FP64 VMR statistics.

A step is one whole estimate (binning, forward FFTs, the 251-rung ladder,
convergence, density) on device-resident samples.  `value` = estimates/s.

Other workloads (--workload): c1 (1-D 512 grid, 10k N(0,1)), c3 (3-D 128^3,
1e6 correlated Gaussian + 10 % uniform), batch (64 heterogeneous datasets,
batched vs plain), serve (Poisson stream through the continuous-batching
server: p50/p99 latency vs the plain per-request path).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "parallelkde.jl_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    import gpu_common as W
    if name == "c1":
        return dict(x=W.gaussian_1d(10000, 1), grid=[(-5.0, 5.0, 512)], B=100,
                    desc="1-D N(0,1), n=10000, 512-point grid over [-5,5], B=100, S=251")
    if name == "c2":
        return dict(x=W.mixture_2d(100000, 2), grid=[(-6.0, 6.0, 256)] * 2, B=100,
                    desc="2-D 3-component mixture, n=100000, 256x256 grid over [-6,6]^2, B=100, S=251")
    if name == "c3":
        return dict(x=W.correlated_3d(1000000, 3), grid=[(-4.0, 4.0, 128)] * 3, B=100,
                    desc="3-D correlated Gaussian (rho=0.5) + 10% uniform, n=1e6, 128^3 grid over "
                         "[-4,4]^3, B=100, S=251")
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(wl, budget_s=20.0):
    """The oracle (FP64 restatement of the reference's :threaded path) on the
    host cores, on a bounded sample: binning + forward FFTs + the first few
    rungs, the ladder time scaled to all 251 rungs."""
    import oracle_ctypes as O
    x = wl["x"]
    lo = [g[0] for g in wl["grid"]]
    hi = [g[1] for g in wl["grid"]]
    n = [g[2] for g in wl["grid"]]
    og = O.make_grid(lo, hi, n)
    threads = min(16, os.cpu_count() or 1)
    # probe one rung to size the sample
    r = O.gradepro(og, x, seed=1, n_bootstraps=wl["B"], threads=threads, max_rungs=1)
    per_rung = max(r.info.seconds_ladder, 1e-6)
    rungs = int(max(1, min(251, (budget_s - r.info.seconds_binning_fft) / per_rung)))
    r = O.gradepro(og, x, seed=1, n_bootstraps=wl["B"], threads=threads, max_rungs=rungs)
    t_est = r.info.seconds_binning_fft + r.info.seconds_ladder * (251.0 / rungs)
    return {"value": 1.0 / t_est, "unit": "estimates/s", "cores": threads, "kind": "port",
            "sample": f"oracle :threaded ({threads} threads) on the full workload: binning + "
                      f"forward FFTs ({r.info.seconds_binning_fft:.2f} s) + {rungs} of 251 rungs "
                      f"({r.info.seconds_ladder:.2f} s), ladder scaled to 251 rungs; "
                      f"projected {t_est:.1f} s per estimate"}


def run_plain(args, wl_name):
    import torch
    import parallelkde_rocm as pk
    from parallelkde_rocm import _lib
    import ctypes as C

    wl = workload(wl_name)
    L = _lib.load()
    ctx = pk.context()
    grid = pk.initialize_grid(wl["grid"])
    params = pk._params(grid.ndims, n_bootstraps=wl["B"], seed=12345, precision=args.precision)
    plan = pk._Plan(grid, params)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(np.ascontiguousarray(wl["x"])).to(dev)
    n = x.shape[0]
    G = int(np.prod(grid.shape))
    dens = torch.empty(G, dtype=torch.float64, device=dev)
    chosen = torch.empty(G, dtype=torch.int16, device=dev)
    stream = torch.cuda.current_stream(dev)
    sptr = C.c_void_p(stream.cuda_stream)

    def step():
        rc = L.pkde_gradepro_estimate_device(plan.handle, x.data_ptr(), n, None, dens.data_ptr(),
                                             chosen.data_ptr(), sptr)
        if rc != 0:
            raise RuntimeError(_lib.last_error())

    log(f"[bench] workload {wl_name}: {wl['desc']}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    if L.pkde_gradepro_check(plan.handle) != 0:
        raise RuntimeError(_lib.last_error())
    ladder_ms = []
    total_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        st = _lib.Stats()
        L.pkde_gradepro_last_stats(plan.handle, C.byref(st))  # waits for this estimate
        ladder_ms.append(st.ms_ladder)
        total_ms.append(st.ms_total)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    S, B = st.n_rungs, wl["B"]
    U = G * S
    # compulsory per-rung re-read of the Hermitian-packed bootstrap spectra
    # (sums and squares) plus the full-sample spectrum: 4 B per real value
    # and grid point in FP32, 8 B in FP64
    per_val = 8.0 if args.precision == "fp64" else 4.0
    ladder_bytes = U * (per_val * (2.0 * B + 1.0))
    lad = float(np.mean(ladder_ms))
    achieved = ladder_bytes / (lad * 1e-3)
    ms_dev = float(np.mean(total_ms))
    out = {
        "metric": "gradepro estimates/s",
        "value": args.steps / wall,
        "unit": "estimates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if args.precision == "fp64"
                  else "f32 spectra/ladder FFT, f64 binning/forward FFT/VMR statistics"),
        "data": "synthetic",
        "config": {"workload": wl["desc"], "grid_points": G, "rungs": S, "bootstraps": B,
                   "samples": n},
        "device_ms_per_estimate": ms_dev,
        "ladder_ms": lad,
        "gp_rungs_per_s": U / (wall / args.steps),
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "ladder (axis-2 inverse pass + fused axis-1 inverse/VMR pass), "
                               "algorithmic bytes U*(2B+1)*(4|8) per estimate over the ladder's "
                               "event-timed duration"},
    }
    plan.close()
    return out, wl


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the :gradepro path has no multi-GPU form in this "
              "build (estimates shard by request; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    out, wl = run_plain(args, args.workload)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle, bounded sample)")
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
