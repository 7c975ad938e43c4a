"""Benchmark of the :gradepro hot path on one MI355X.

Default workload (--workload batch): the 64-dataset batched mix, the largest
single-GPU configuration of the benchmark set -- dataset j (seed 1000 + j) has
a dimension uniform in {1, 2, 3}, a sample count log-uniform in [1e3, 2e5],
N(0,I) or 2-component mixture draws, and a grid of 1024 (1-D), 128^2 (2-D) or
64^3 (3-D) points over [-6, 6] per axis; B = 100 bootstraps, S = 251 rungs,
FP64 throughout.  A step is one batched submission of all 64 estimates
(pkde_gradepro_estimate_batch_device: binning, forward transforms, the
251-rung ladder, convergence, density), samples already in device memory;
`value` = estimates/s.  The plain form -- one request per launch sequence,
each estimate synchronised before the next (pkde_gradepro_estimate_device) --
is timed on the same data, overall and per dimension class.

Other workloads (--workload): c1 (1-D 512 grid, 1e4 N(0,1)), c2 (2-D 256^2,
1e5 3-component mixture), c3 (3-D 128^3, 1e6 correlated Gaussian + 10 %
uniform), 3d64 (one 64^3 dataset of the batched kind), serve (Poisson stream
through the continuous-batching server at a rate above the plain path's
capacity: p50/p99 latency and estimates/s against the plain per-request
path), copy (a device copy of known size, for the HBM counter calibration).

Roofline: A = 8 B + 4 = 804 algorithmic bytes per grid-point * rung at
B = 100 (the per-rung re-read of Hermitian-packed FP32 bootstrap spectra, sums
and squares, plus the full-sample spectrum).  `roofline.achieved` divides
U * A (U = G * S grid-point * rungs) by the measured time: the whole batch's
wall time for the batch, the event-timed ladder for single workloads; the
FP64 path moves twice these bytes per value, `frac_fp64_bytes` says so.

The CPU baseline is the FP64 oracle (a C restatement of the reference's
:threaded path, test infrastructure) on the host cores, on a bounded sample;
the same oracle runs give the accuracy figures (relative L1 / Linf density
error and chosen-rung agreement against the GPU results of the timed run,
same samples and bootstrap draws).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batch]
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
CPU_THREADS = 16   # host cores of a one-GPU box


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(G, S, B=100):
    """Algorithmic ladder bytes of one estimate: U * (8 B + 4)."""
    return float(G) * S * (8.0 * B + 4.0)


def measured_traffic(name):
    """HBM bytes per step of the ladder kernels (copy-calibrated FETCH_SIZE +
    WRITE_SIZE) from the committed rocprofv3 counter summary of the same
    workload, or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_step:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


# ------------------------------------------------------------ single runs
def workload(name):
    import gpu_common as W
    if name == "c1":
        return dict(x=W.gaussian_1d(10000, 1), grid=[(-5.0, 5.0, 512)], B=100, seed=1,
                    desc="1-D N(0,1), n=10000, 512-point grid over [-5,5], B=100, S=251")
    if name == "c2":
        return dict(x=W.mixture_2d(100000, 2), grid=[(-6.0, 6.0, 256)] * 2, B=100, seed=2,
                    desc="2-D 3-component mixture, n=100000, 256x256 grid over [-6,6]^2, B=100, "
                         "S=251")
    if name == "c3":
        return dict(x=W.correlated_3d(1000000, 3), grid=[(-4.0, 4.0, 128)] * 3, B=100, seed=3,
                    desc="3-D correlated Gaussian (rho=0.5) + 10% uniform, n=1e6, 128^3 grid over "
                         "[-4,4]^3, B=100, S=251")
    if name == "3d64":
        return dict(x=W.correlated_3d(200000, 4, lo=-6.0, hi=6.0), grid=[(-6.0, 6.0, 64)] * 3,
                    B=100, seed=4, desc="3-D correlated Gaussian (rho=0.5) + 10% uniform, n=2e5, "
                                        "64^3 grid over [-6,6]^3, B=100, S=251")
    raise SystemExit(f"unknown workload {name}")


def oracle_estimate_time(x, ranges, seed, B=100, budget_s=10.0):
    """Seconds the oracle (:threaded, CPU_THREADS threads) takes for one whole
    estimate of x on `ranges`: binning and forward transforms measured, the
    ladder measured on as many rungs as fit `budget_s` and scaled to 251.
    Returns (seconds, rungs run, result of the run)."""
    import oracle_ctypes as O
    og = O.make_grid([r[0] for r in ranges], [r[1] for r in ranges], [r[2] for r in ranges])
    r = O.gradepro(og, x, seed=seed, n_bootstraps=B, threads=CPU_THREADS, max_rungs=1)
    per_rung = max(r.info.seconds_ladder, 1e-6)
    rungs = int(max(1, min(251, (budget_s - r.info.seconds_binning_fft) / per_rung)))
    if rungs > 1:
        r = O.gradepro(og, x, seed=seed, n_bootstraps=B, threads=CPU_THREADS, max_rungs=rungs)
    t = r.info.seconds_binning_fft + r.info.seconds_ladder * (251.0 / rungs)
    return t, rungs, r


def accuracy_of(pairs):
    """pairs: [(gpu density, gpu chosen, oracle result)] of whole ladders."""
    import gpu_common as W
    ms = [W.compare(d, c, r.density, r.chosen, 251) for d, c, r in pairs]
    if not ms:
        return None
    l1 = np.array([m["rel_l1"] for m in ms])
    out = {"estimates_compared": len(ms),
           "rel_l1_max": float(l1.max()),
           "rel_linf_max": max(m["rel_linf"] for m in ms),
           "chosen_agree_all_min": min(m["agree_all"] for m in ms),
           "chosen_agree_significant_min": min(m["agree_sig"] for m in ms),
           "significant": "points whose oracle density is >= 1e-3 of its maximum",
           "oracle": "FP64 oracle on the same samples and bootstrap draws"}
    if len(ms) > 1:
        sig = np.array([m["agree_sig"] for m in ms])
        out.update({"rel_l1_median": float(np.median(l1)),
                    "estimates_within_1e-5_rel_l1": int((l1 <= 1e-5).sum()),
                    "estimates_with_99.9pct_significant_agreement": int((sig >= 0.999).sum()),
                    "note": "the largest errors are datasets of 1e3-4e3 samples, where the "
                            "oracle's own result moves as far under FP64 rounding alone"})
    return out


def run_plain(args, wl_name):
    import ctypes as C
    import torch
    import parallelkde_rocm as pk
    from parallelkde_rocm import _lib

    wl = workload(wl_name)
    L = _lib.load()
    torch.cuda.init()
    pk.context()
    grid = pk.initialize_grid(wl["grid"])
    params = pk._params(grid.ndims, n_bootstraps=wl["B"], seed=wl["seed"], precision=args.precision)
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
    ladder_ms, total_ms = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        st = _lib.Stats()
        L.pkde_gradepro_last_stats(plan.handle, C.byref(st))  # waits for this estimate
        ladder_ms.append(st.ms_ladder)
        total_ms.append(st.ms_total)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if L.pkde_gradepro_check(plan.handle) != 0:
        raise RuntimeError(_lib.last_error())
    S = st.n_rungs
    U = G * S
    lad = float(np.mean(ladder_ms))
    bytes_alg = alg_bytes(G, S, wl["B"])
    achieved = bytes_alg / (lad * 1e-3)
    per_val = 8.0 if args.precision == "fp64" else 4.0
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
        "config": {"workload": wl["desc"], "grid_points": G, "rungs": S, "bootstraps": wl["B"],
                   "samples": n},
        "device_ms_per_estimate": float(np.mean(total_ms)),
        "ladder_ms": lad,
        "gp_rungs_per_s": U / (wall / args.steps),
        "roofline": {
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "frac_fp64_bytes": achieved * per_val / 4.0 / HBM_PEAK,
            "traffic": measured_traffic(f"{wl_name}_hbm_traffic.txt"),
            "traffic_unit": "HBM bytes per estimate, ladder kernels",
            "algorithmic_bytes": bytes_alg,
            "kernel": "the ladder (axis passes + row/slab kernels of every rung chunk), between "
                      "two HIP events on its stream; algorithmic bytes U*(8B+4)"},
    }
    gpu = (dens.cpu().numpy(), chosen.cpu().numpy())
    plan.close()
    return out, wl, gpu


def cpu_leg_single(wl, gpu, budget_s):
    """Oracle on the same samples and draws: the whole estimate when it fits
    the budget (then also the accuracy of the GPU result), else a bounded
    ladder scaled to 251 rungs."""
    t, rungs, r = oracle_estimate_time(wl["x"], wl["grid"], wl["seed"], wl["B"], budget_s)
    base = {"value": 1.0 / t, "unit": "estimates/s", "cores": CPU_THREADS, "kind": "port",
            "sample": f"oracle :threaded ({CPU_THREADS} threads), same samples and draws: binning + "
                      f"forward FFTs ({r.info.seconds_binning_fft:.2f} s) + {rungs} of 251 rungs "
                      f"({r.info.seconds_ladder:.2f} s)"
                      + (", ladder scaled to 251 rungs" if rungs < 251 else "")
                      + f"; {t:.2f} s per estimate"}
    acc = accuracy_of([(gpu[0], gpu[1], r)]) if rungs >= 251 else None
    return base, acc


# ------------------------------------------------------------------ batch
def run_batch(args):
    import ctypes as C
    import torch
    import gpu_common as W
    import parallelkde_rocm as pk
    from parallelkde_rocm import _lib

    L = _lib.load()
    torch.cuda.init()
    ctx = pk.context()
    dev = torch.device("cuda:0")
    items = W.batch_mix_items()
    shapes = {}
    xs, dens, chos, grids = [], [], [], []
    for x, ranges, seed in items:
        g = shapes.setdefault(tuple(ranges), pk.initialize_grid(ranges))
        grids.append(g)
        xs.append(torch.from_numpy(np.ascontiguousarray(x)).to(dev))
        G = int(np.prod(g.shape))
        dens.append(torch.empty(G, dtype=torch.float64, device=dev))
        chos.append(torch.empty(G, dtype=torch.int16, device=dev))
    plain_seed = 0x5EED

    def requests(idx, seeded=True):
        reqs = (_lib.Request * len(idx))()
        for k, i in enumerate(idx):
            g, r = grids[i], reqs[k]
            r.ndims = g.ndims
            for a in range(g.ndims):
                r.n_per_axis[a], r.lo[a], r.hi[a], r.h[a] = g.n[a], g.lo[a], g.hi[a], g.h[a]
            r.samples = xs[i].data_ptr()
            r.n = xs[i].shape[0]
            r.params = pk._params(g.ndims, seed=items[i][2] if seeded else plain_seed,
                                  precision=args.precision)
            r.density_out = dens[i].data_ptr()
            r.chosen_out = chos[i].data_ptr()
        return reqs

    def batch(reqs, count):
        status = np.zeros(count, dtype=np.int32)
        rc = L.pkde_gradepro_estimate_batch_device(ctx.handle, count, reqs, status.ctypes.data)
        if rc != 0 or status.any():
            raise RuntimeError(f"batch: {_lib.last_error()} {status}")

    # plain form: one plan per grid shape (the default seed), one request at
    # a time on the context stream, each synchronised before the next
    plans = {key: pk._Plan(g, pk._params(g.ndims, seed=plain_seed, precision=args.precision))
             for key, g in shapes.items()}
    pd = [torch.empty_like(d) for d in dens]
    pc = [torch.empty_like(c) for c in chos]
    ladder_ms = {1: [], 2: [], 3: []}

    def plain(idx, record=False):
        for i in idx:
            pl = plans[tuple(items[i][1])]
            rc = L.pkde_gradepro_estimate_device(pl.handle, xs[i].data_ptr(), xs[i].shape[0], None,
                                                 pd[i].data_ptr(), pc[i].data_ptr(), None)
            if rc != 0:
                raise RuntimeError(_lib.last_error())
            if record:
                st = _lib.Stats()
                L.pkde_gradepro_last_stats(pl.handle, C.byref(st))
                ladder_ms[grids[i].ndims].append(st.ms_ladder)
            torch.cuda.synchronize()

    everything = list(range(len(items)))
    by_dim = {d: [i for i in everything if grids[i].ndims == d] for d in (1, 2, 3)}
    reqs_all = requests(everything)
    log(f"[bench] batch: {len(items)} datasets ({len(by_dim[1])} 1-D / {len(by_dim[2])} 2-D / "
        f"{len(by_dim[3])} 3-D); warmup")
    for _ in range(max(1, args.warmup)):
        batch(reqs_all, len(items))
        if not args.batch_only:
            plain(everything)
    torch.cuda.synchronize()

    def timed(fn, reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    t_batch = timed(lambda: batch(reqs_all, len(items)), args.steps)
    log(f"[bench] batch: {1e3 * t_batch:.1f} ms per batch")
    if args.batch_only:
        # profiling runs: the batched submissions alone (warmup + steps)
        for p in plans.values():
            p.close()
        pk.release_batch_memory()
        return {"workload": "batch only", "batches": max(1, args.warmup) + args.steps,
                "ms_per_batch": 1e3 * t_batch, "estimates_per_s": len(items) / t_batch}, items, None
    gpu_out = [(d.cpu().numpy(), c.cpu().numpy()) for d, c in zip(dens, chos)]
    t_plain = timed(lambda: plain(everything), args.steps)
    plain(everything, record=True)
    log(f"[bench] plain: {1e3 * t_plain:.1f} ms per 64 requests")
    per_class = {}
    for d in (1, 2, 3):
        idx = by_dim[d]
        if not idx:
            continue
        rq = requests(idx)
        batch(rq, len(idx))
        tb = timed(lambda: batch(rq, len(idx)), args.steps)
        tp = timed(lambda: plain(idx), args.steps)
        per_class[f"{d}-D"] = {"requests": len(idx), "batched_estimates_per_s": len(idx) / tb,
                               "plain_estimates_per_s": len(idx) / tp,
                               "batched_over_plain": tp / tb}
    # the batch with the plain plans' seed equals the plain form bit for bit
    batch(requests(everything, seeded=False), len(items))
    plain(everything)
    same = all(torch.equal(dens[i], pd[i]) and torch.equal(chos[i], pc[i]) for i in everything)
    S = 251
    total_alg = sum(alg_bytes(int(np.prod(g.shape)), S) for g in grids)
    achieved = total_alg / t_batch
    G3 = 64 ** 3
    lad3 = float(np.mean(ladder_ms[3])) if ladder_ms[3] else None
    dom = None
    if lad3:
        a3 = alg_bytes(G3, S) / (lad3 * 1e-3)
        dom = {"kernel": "64^3 ladder (k_axis_ladder<double,6> + k_slab3<6,6> per rung chunk), "
                         "event-timed on its stream in the plain runs",
               "ladder_ms": lad3, "achieved": a3 / 1e9, "frac": a3 / HBM_PEAK,
               "frac_fp64_bytes": 2.0 * a3 / HBM_PEAK}
    out = {
        "metric": "gradepro estimates/s",
        "value": len(items) / t_batch,
        "unit": "estimates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(1, args.warmup),
        "ms_per_step": 1e3 * t_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 ladder",
        "data": "synthetic",
        "config": {"workload": "batched mix: 64 datasets in one submission, d in {1,2,3} "
                               f"({len(by_dim[1])}/{len(by_dim[2])}/{len(by_dim[3])}), n log-uniform "
                               "1e3..2e5, N(0,I) or 2-component mixture, grids 1024 / 128^2 / 64^3 "
                               "over [-6,6]^d, B=100, S=251, samples resident in device memory",
                   "datasets": len(items), "bootstraps": 100, "rungs": S},
        "plain_estimates_per_s": len(items) / t_plain,
        "batched_over_plain": t_plain / t_batch,
        "per_class": per_class,
        "results_identical_to_plain": bool(same),
        "roofline": {
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "frac_fp64_bytes": 2.0 * achieved / HBM_PEAK,
            "traffic": measured_traffic("c4_batch_hbm_traffic.txt"),
            "traffic_unit": "HBM bytes per batch, every kernel",
            "algorithmic_bytes": total_alg,
            "kernel": "whole batch: algorithmic ladder bytes U*(8B+4) summed over the 64 "
                      "estimates over the batch's wall time (binning, transforms, convergence "
                      "included)",
            "dominant": dom},
    }
    for p in plans.values():
        p.close()
    pk.release_batch_memory()
    return out, items, gpu_out


def cpu_leg_batch(items, gpu_out, budget_s):
    """Oracle time for the whole mix from one dataset per dimension class
    (bounded ladders scaled to 251 rungs), and the accuracy of the 1-D items
    and the first 2-D items against whole oracle ladders."""
    import oracle_ctypes as O
    t_start = time.time()
    per_class, counts = {}, {1: 0, 2: 0, 3: 0}
    for x, ranges, seed in items:
        counts[len(ranges)] += 1
    for d in (1, 2, 3):
        first = next(((x, r, s) for x, r, s in items if len(r) == d), None)
        if first is None:
            continue
        t, rungs, _ = oracle_estimate_time(first[0], first[1], first[2], 100, budget_s / 4)
        per_class[d] = (t, rungs, first[0].shape[0])
    total = sum(per_class[d][0] * counts[d] for d in per_class)
    pairs = []
    for j, (x, ranges, seed) in enumerate(items):
        if len(ranges) == 3 or time.time() - t_start > 2.0 * budget_s:
            continue
        og = O.make_grid([r[0] for r in ranges], [r[1] for r in ranges], [r[2] for r in ranges])
        r = O.gradepro(og, x, seed=seed, threads=CPU_THREADS)
        pairs.append((gpu_out[j][0], gpu_out[j][1], r))
    sample = "; ".join(f"{d}-D item n={per_class[d][2]}: {per_class[d][0]:.2f} s "
                       f"({per_class[d][1]} of 251 rungs run)" for d in per_class)
    base = {"value": len(items) / total, "unit": "estimates/s", "cores": CPU_THREADS, "kind": "port",
            "sample": f"oracle :threaded ({CPU_THREADS} threads) on the first dataset of each "
                      f"class, ladders scaled to 251 rungs, times the class counts "
                      f"({counts[1]}/{counts[2]}/{counts[3]}): {sample}; projected "
                      f"{total:.1f} s per batch of {len(items)}"}
    acc = accuracy_of(pairs)
    if acc:
        acc["items"] = "1-D and 2-D items of the batch, whole ladders, as many as fit the budget"
    return base, acc


# ------------------------------------------------------------------ serve
def run_serve(args):
    """Poisson arrivals through the continuous-batching server against the
    plain per-request path (FIFO, one estimate at a time) on the same arrival
    times, at two offered rates: 1.5x the plain path's capacity (1 / mean
    service time of the requests alone), where both saturate and the
    throughputs compare, and 0.6x, where the latencies compare."""
    import threading
    import gpu_common as W
    import parallelkde_rocm as pk
    count = args.serve_requests
    reqs = W.stream_requests(count)
    grids = [pk.initialize_grid(r) for _, r, _ in reqs]
    plans = {}
    for (x, r, s), g in zip(reqs, grids):
        if tuple(g.n) not in plans:
            plans[tuple(g.n)] = pk._Plan(g, pk._params(g.ndims, precision=args.precision))
    # warm every shape, then the mean service time of the plain path
    for (x, r, s), g in zip(reqs, grids):
        plans[tuple(g.n)].estimate(np.ascontiguousarray(x))
    t0 = time.perf_counter()
    for (x, r, s), g in zip(reqs, grids):
        plans[tuple(g.n)].estimate(np.ascontiguousarray(x))
    service = (time.perf_counter() - t0) / count
    srv = pk.Server(max_inflight=args.serve_inflight)
    seen = set()
    warm = []
    for (x, r, s), g in zip(reqs, grids):  # warm the server's pool per shape, all slots at once
        if tuple(g.n) not in seen:
            seen.add(tuple(g.n))
            warm.append(srv.submit(x, g, precision=args.precision))
    for t in warm:
        srv.wait(t)
    q = lambda v, f: float(np.quantile(np.array(v) * 1e3, f))

    def at_rate(rate, seed):
        rng = np.random.default_rng(seed)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
        lat_plain = []
        t0 = time.perf_counter()
        for (x, r, s), g, ta in zip(reqs, grids, arrivals):
            now = time.perf_counter() - t0
            if now < ta:
                time.sleep(ta - now)
            plans[tuple(g.n)].estimate(np.ascontiguousarray(x))
            lat_plain.append(time.perf_counter() - t0 - ta)
        span_plain = time.perf_counter() - t0
        lat_srv = [None] * count
        tickets = [None] * count
        done_ev = threading.Event()
        t0 = time.perf_counter()

        def waiter():
            # poll every outstanding ticket: a request's completion time is
            # recorded when it completes, whatever the order.  Every 2 ms: a
            # poll takes the server's lock, and polling much faster measurably
            # slows its worker thread (latencies here are 30 ms to 2 s)
            left = set(range(count))
            while left:
                for i in sorted(left):
                    if tickets[i] is None:
                        continue
                    if srv.poll(tickets[i]) is not None:
                        lat_srv[i] = time.perf_counter() - t0 - arrivals[i]
                        left.discard(i)
                time.sleep(0.002)
            done_ev.set()

        th = threading.Thread(target=waiter)
        th.start()
        for i, ((x, r, s), g, ta) in enumerate(zip(reqs, grids, arrivals)):
            now = time.perf_counter() - t0
            if now < ta:
                time.sleep(ta - now)
            tickets[i] = srv.submit(x, g, precision=args.precision)
        done_ev.wait()
        th.join()
        span_srv = time.perf_counter() - t0
        return {"offered_per_s": rate, "served_estimates_per_s": count / span_srv,
                "plain_estimates_per_s": count / span_plain,
                "served_over_plain": span_plain / span_srv,
                "p50_ms": q(lat_srv, 0.5), "p99_ms": q(lat_srv, 0.99),
                "plain_p50_ms": q(lat_plain, 0.5), "plain_p99_ms": q(lat_plain, 0.99)}

    cap = 1.0 / service
    rate = args.serve_rate if args.serve_rate > 0 else 1.5 * cap
    log(f"[bench] serve: {count} requests, plain capacity {cap:.2f}/s, offered {rate:.2f}/s")
    sat = at_rate(rate, 99)
    log(f"[bench] serve: offered {0.6 * cap:.2f}/s")
    mod = at_rate(0.6 * cap, 98)
    st = srv.stats()
    srv.close()
    for p in plans.values():
        p.close()
    dims = {d: sum(1 for _, r, _ in reqs if len(r) == d) for d in (1, 2, 3)}
    return {
        "metric": "gradepro estimates/s (served Poisson stream)",
        "value": sat["served_estimates_per_s"], "unit": "estimates/s", "n_gpus": 1,
        "steps": count, "warmup": len(seen), "ms_per_step": 1e3 / sat["served_estimates_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 ladder", "data": "synthetic",
        "config": {"workload": f"{count} requests, Poisson at {rate:.2f}/s (1.5x the plain "
                               f"path's capacity {cap:.2f}/s), d in {{1,2,3}} "
                               f"({dims[1]}/{dims[2]}/{dims[3]}), n log-uniform 1e3..1e6, grids "
                               "256-1024 / 64-256^2 / 32-128^3 (powers of two), B=100, S=251, host "
                               "buffers in and out", "max_inflight": args.serve_inflight},
        "p50_ms": sat["p50_ms"], "p99_ms": sat["p99_ms"],
        "plain_estimates_per_s": sat["plain_estimates_per_s"],
        "plain_p50_ms": sat["plain_p50_ms"], "plain_p99_ms": sat["plain_p99_ms"],
        "served_over_plain": sat["served_over_plain"],
        "saturated": sat, "moderate_load": mod,
        "server_bytes_in_use": st["bytes_in_use"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="batch workload: time the batched submissions only (profiling runs)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--serve-rate", type=float, default=0.0)
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
    if args.workload == "serve":
        print(json.dumps(run_serve(args)), flush=True)
        return
    if args.workload == "batch":
        out, items, gpu_out = run_batch(args)
        if not args.no_cpu_baseline and not args.batch_only:
            log("[bench] cpu baseline and accuracy (oracle, bounded sample)")
            out["cpu_baseline"], out["accuracy"] = cpu_leg_batch(items, gpu_out, args.cpu_budget)
        print(json.dumps(out), flush=True)
        return
    out, wl, gpu = run_plain(args, args.workload)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline and accuracy (oracle, bounded sample)")
        out["cpu_baseline"], out["accuracy"] = cpu_leg_single(wl, gpu, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
