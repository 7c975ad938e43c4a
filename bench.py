"""Benchmark of the :gradepro hot path on one MI355X.

Default workload (the 2-D configuration the headline metric is quoted on):
  100,000 samples from an equal-weight 3-component Gaussian mixture (means
  (-2,0), (2,0), (0,2.5), isotropic sd 0.6, seed 2), 256 x 256 grid over
  [-6,6]^2, B = 100 bootstraps, 251 bandwidth rungs, FP64 throughout (the
  parity-grade mode; --precision fp32 keeps spectra and ladder FFTs in FP32).

A step is one whole estimate (binning, forward FFTs, the 251-rung ladder,
convergence, density) on device-resident samples.  `value` = estimates/s.

Other workloads (--workload): c1 (1-D 512 grid, 10k N(0,1)), c3 (3-D 128^3,
1e6 correlated Gaussian + 10 % uniform), 3d64 (a 64^3 dataset of the kind the
batched mix holds: 2e5 correlated Gaussian + 10 % uniform over [-6,6]^3),
batch (64 heterogeneous datasets,
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
    if name == "3d64":
        return dict(x=W.correlated_3d(200000, 4, lo=-6.0, hi=6.0), grid=[(-6.0, 6.0, 64)] * 3,
                    B=100, desc="3-D correlated Gaussian (rho=0.5) + 10% uniform, n=2e5, 64^3 grid "
                                "over [-6,6]^3, B=100, S=251")
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
    if args.warmup and L.pkde_gradepro_check(plan.handle) != 0:
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
    if L.pkde_gradepro_check(plan.handle) != 0:
        raise RuntimeError(_lib.last_error())
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
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": _measured_traffic(wl_name, args.precision),
                     "traffic_unit": "bytes per estimate (ladder kernels)",
                     "algorithmic_bytes": ladder_bytes,
                     "kernel": "ladder (axis-2 inverse pass + fused axis-1 inverse/VMR pass), "
                               "algorithmic bytes U*(2B+1)*(4|8) per estimate over the ladder's "
                               "event-timed duration; traffic = calibrated FETCH_SIZE+WRITE_SIZE "
                               "bytes of the same kernels per estimate"},
    }
    plan.close()
    return out, wl


def _measured_traffic(wl_name, precision):
    """HBM bytes per estimate of the ladder kernels, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the same workload (copy-kernel
    calibrated); None when no such summary exists."""
    files = {("c2", "fp64"): "c2_2d256_hbm_traffic.txt"}
    name = files.get((wl_name, precision))
    if name is None:
        return None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("ladder_bytes_per_estimate:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def _box_draws(rng, n, d, lo, hi, kind, margin):
    """N(0,I) or the mixture 1/2 N(-1.5*1, 0.5^2 I) + 1/2 N(+1.5*1, I), draws
    outside (lo, hi - margin) redrawn."""
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


def c4_items():
    """64 independent datasets: dimension uniform in {1,2,3}, sample counts
    log-uniform in [1e3, 2e5], N(0,I) or 2-component mixture draws; grids
    1024 (1-D), 128^2 (2-D), 64^3 (3-D) over [-6,6] per axis."""
    items = []
    for j in range(64):
        rng = np.random.default_rng(1000 + j)
        d = int(rng.integers(1, 4))
        n = int(10 ** rng.uniform(3, math.log10(2e5)))
        kind = "normal" if rng.random() < 0.5 else "mixture"
        pts = {1: 1024, 2: 128, 3: 64}[d]
        x = _box_draws(rng, n, d, -6.0, 6.0, kind, 12.0 / (pts - 1) + 1e-9)
        items.append((x, [(-6.0, 6.0, pts)] * d, {}))
    return items


def run_batch(args):
    """Batched (one submission, concurrent stream lanes) vs plain (one request
    per synchronous call, plans reused per shape) on the 64-dataset mix."""
    import parallelkde_rocm as pk
    items = c4_items()
    grids = {}
    for x, g, kw in items:
        grids.setdefault(tuple(g), pk.initialize_grid(g))
    req = [(x, grids[tuple(g)], dict(kw, precision=args.precision)) for x, g, kw in items]
    plans = {}

    def plain_all():
        out = []
        for x, grid, kw in req:
            key = tuple(grid.n)
            if key not in plans:
                plans[key] = pk._Plan(grid, pk._params(grid.ndims, precision=args.precision))
            out.append(plans[key].estimate(np.ascontiguousarray(x)).density)
        return out

    log("[bench] batch: warmup")
    plain_all()
    # the ladder launch shape of each of the three grids, tuned on the device
    # (pkde_gradepro_autotune); the batch's lane plans take it from the context
    tuned = {}
    for key, p in plans.items():
        best, default = p.autotune(reps=2)
        tuned["x".join(map(str, key))] = {"ladder_ms": round(best, 3),
                                          "default_shape_ladder_ms": round(default, 3)}
    log(f"[bench] batch: autotuned {tuned}")
    plain_all()
    pk.estimate_batch(req)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dp = plain_all()
    t_plain = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        db = pk.estimate_batch(req)
    t_batch = (time.perf_counter() - t0) / args.steps
    same = all(np.array_equal(a, b) for a, b in zip(dp, db))
    per_val = 8.0 if args.precision == "fp64" else 4.0
    alg = sum(np.prod(g.shape) * 251 * per_val * 201 for _, g, _ in req)
    counts = {d: sum(1 for _, g, _ in req if g.ndims == d) for d in (1, 2, 3)}
    out = {
        "metric": "gradepro estimates/s (batched 64-dataset mix)",
        "value": len(req) / t_batch, "unit": "estimates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": 1e3 * t_batch, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32 ladder",
        "data": "synthetic",
        "config": {"workload": "64 datasets, d in {1,2,3} "
                               f"({counts[1]}/{counts[2]}/{counts[3]}), n log-uniform 1e3..2e5, "
                               "grids 1024 / 128^2 / 64^3 over [-6,6]^d, B=100, S=251"},
        "plain_estimates_per_s": len(req) / t_plain,
        "batched_over_plain": t_plain / t_batch,
        "results_identical_to_plain": bool(same),
        "autotune": tuned,
        "roofline": {"bound": "hbm", "achieved": alg / t_batch / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": alg / t_batch / HBM_PEAK, "traffic": None,
                     "kernel": "whole batch, algorithmic ladder bytes U*(2B+1)*8 summed over "
                               "requests over the batch wall time"},
    }
    for p in plans.values():
        p.close()
    return out


def c5_requests(count, seed=7):
    """Request r (seed r): d uniform in {1,2,3}; n log-uniform in [1e3, 1e6];
    per-axis points log-uniform in 256-1024 (1-D), 64-256 (2-D), 32-128 (3-D),
    rounded to the nearest power of two; N(0,I) draws on [-6,6]^d."""
    reqs = []
    for r in range(count):
        rng = np.random.default_rng(seed * 100000 + r)
        d = int(rng.integers(1, 4))
        n = int(10 ** rng.uniform(3, 6))
        lo_p, hi_p = {1: (256, 1024), 2: (64, 256), 3: (32, 128)}[d]
        pts = int(2 ** round(rng.uniform(math.log2(lo_p), math.log2(hi_p))))
        x = _box_draws(rng, n, d, -6.0, 6.0, "normal", 12.0 / (pts - 1) + 1e-9)
        reqs.append((x, [(-6.0, 6.0, pts)] * d))
    return reqs


def run_serve(args):
    """Poisson arrivals at a fixed offered rate through the continuous-batching
    server vs the plain per-request path (FIFO, one estimate at a time) on the
    same arrival times."""
    import threading
    import parallelkde_rocm as pk
    count, rate = args.serve_requests, args.serve_rate
    reqs = c5_requests(count)
    rng = np.random.default_rng(99)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    grids = [pk.initialize_grid(g) for _, g in reqs]
    # warm both paths on every distinct shape (plan creation, code objects)
    shapes = {}
    for (x, g), grid in zip(reqs, grids):
        shapes.setdefault(tuple(grid.n), (x[:2000], grid))
    plans = {}
    for key, (x, grid) in shapes.items():
        plans[key] = pk._Plan(grid, pk._params(grid.ndims, precision=args.precision))
        plans[key].estimate(np.ascontiguousarray(x))
    log(f"[bench] serve: {count} requests at {rate}/s, plain path")
    # plain: FIFO, one request at a time
    lat_plain = []
    t0 = time.perf_counter()
    for (x, g), grid, ta in zip(reqs, grids, arrivals):
        now = time.perf_counter() - t0
        if now < ta:
            time.sleep(ta - now)
        plans[tuple(grid.n)].estimate(np.ascontiguousarray(x))
        lat_plain.append(time.perf_counter() - t0 - ta)
    span_plain = time.perf_counter() - t0
    for p in plans.values():
        p.close()
    log("[bench] serve: server path")
    srv = pk.Server(max_inflight=args.serve_inflight)
    for key, (x, grid) in shapes.items():
        srv.wait(srv.submit(x, grid, precision=args.precision))
    lat_srv = [None] * count
    done_ev = threading.Event()
    tickets = [None] * count
    t0 = time.perf_counter()

    def waiter():
        for i in range(count):
            while tickets[i] is None:
                time.sleep(0.0005)
            srv.wait(tickets[i])
            lat_srv[i] = time.perf_counter() - t0 - arrivals[i]
        done_ev.set()

    th = threading.Thread(target=waiter)
    th.start()
    for i, ((x, g), grid, ta) in enumerate(zip(reqs, grids, arrivals)):
        now = time.perf_counter() - t0
        if now < ta:
            time.sleep(ta - now)
        tickets[i] = srv.submit(x, grid, precision=args.precision)
    done_ev.wait()
    th.join()
    span_srv = time.perf_counter() - t0
    srv.close()
    q = lambda v, f: float(np.quantile(np.array(v) * 1e3, f))
    dims = {d: sum(1 for _, g in reqs if len(g) == d) for d in (1, 2, 3)}
    return {
        "metric": "gradepro estimates/s (served Poisson stream)",
        "value": count / span_srv, "unit": "estimates/s", "n_gpus": 1, "steps": count,
        "warmup": len(shapes), "ms_per_step": 1e3 * span_srv / count, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 ladder", "data": "synthetic",
        "config": {"workload": f"{count} requests, Poisson {rate}/s, d in {{1,2,3}} "
                               f"({dims[1]}/{dims[2]}/{dims[3]}), n log-uniform 1e3..1e6, "
                               "grids 256-1024 / 64-256^2 / 32-128^3 (powers of two), B=100, S=251",
                   "max_inflight": args.serve_inflight},
        "p50_ms": q(lat_srv, 0.5), "p99_ms": q(lat_srv, 0.99),
        "plain_estimates_per_s": count / span_plain,
        "plain_p50_ms": q(lat_plain, 0.5), "plain_p99_ms": q(lat_plain, 0.99),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=3.0)
    ap.add_argument("--serve-inflight", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the :gradepro path has no multi-GPU form here "
              "(independent estimates shard by request, one process per GPU).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        # calibration for the HBM byte counters: a device copy of known size
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"workload": "copy", "bytes_read": a.numel() * 4,
                          "bytes_written": b.numel() * 4}), flush=True)
        return
    if args.workload == "batch":
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(run_serve(args)), flush=True)
        return
    out, wl = run_plain(args, args.workload)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle, bounded sample)")
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
