#!/usr/bin/env python3
"""Benchmark: Mvis/s of the gridder on an 8192^2 grid, support 64, 10M
radially clustered visibilities (FP32-tiled vs FP64 plain scatter, with
per-step p50/p99 latency), one JSON line on stdout.

A step is one full gridding pass over the 10M-visibility batch resident in
HBM: grid zeroing, bin sort (tiled paths) and the gridding kernels.  The
timed region is bracketed by torch.cuda.synchronize() on both sides; each
step is also timed with HIP events for the latency percentiles.

Inputs are synthetic, drawn here with a counter-based SplitMix64 stream
(seed 1541618): angle U[0, 2pi), radius Exp(mean 800 px) around the grid
centre, redrawn until the 64 x 64 window lies on the grid, uv rounded to
float32-representable values, values complex N(0,1) (Box-Muller).  The GCF
is the separable Gaussian (sigma = S/6) x exp(0.05 i r^2) table.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ska-gpu-grid_amd"))

METRIC = "Mvis/s at 8192² grid, support 64: FP32-tiled vs FP64 plain; p99 latency"
PEAK_FP64_TF = 78.6   # MI355X FP64 vector/matrix, dense
PEAK_FP32_TF = 157.3  # MI355X FP32 vector (packed), dense
PEAK_HBM_GBS = 8000.0

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_u01(seed, counter):
    """Uniform [0,1) doubles from a counter-based SplitMix64 stream."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + _G * (counter.astype(np.uint64) + np.uint64(1))
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def gen_clustered(npts, N, S, mean_r=800.0, seed=1541618):
    h = S // 2
    uv = np.empty((npts, 2), np.float64)
    todo = np.arange(npts)
    rnd = 0
    while todo.size:
        c = (np.int64(rnd) * npts * 2 + np.arange(todo.size) * 2).astype(np.uint64)
        th = 2 * math.pi * splitmix_u01(seed, c)
        r = -mean_r * np.log1p(-splitmix_u01(seed, c + np.uint64(1)))
        u = (N / 2 + r * np.cos(th)).astype(np.float32).astype(np.float64)
        v = (N / 2 + r * np.sin(th)).astype(np.float32).astype(np.float64)
        fu, fv = np.floor(u), np.floor(v)
        ok = (fu - h + 1 >= 0) & (fu + h <= N - 1) & (fv - h + 1 >= 0) & (fv + h <= N - 1)
        uv[todo[ok], 0] = u[ok]
        uv[todo[ok], 1] = v[ok]
        todo = todo[~ok]
        rnd += 1
    c = np.arange(npts, dtype=np.uint64) * np.uint64(2) + np.uint64(1 << 40)
    u1 = 1.0 - splitmix_u01(seed, c)
    u2 = splitmix_u01(seed, c + np.uint64(1))
    rr = np.sqrt(-2.0 * np.log(u1))
    vals = np.stack([rr * np.cos(2 * math.pi * u2), rr * np.sin(2 * math.pi * u2)], 1)
    return uv, vals


def gen_uniform(npts, N, S, seed=1541618):
    """Uniform uv over the grid interior (every window on the grid)."""
    h = S // 2
    c = np.arange(npts, dtype=np.uint64) * np.uint64(4)
    lo, hi = h - 1.0, N - h - 1.0
    u = (lo + (hi - lo) * splitmix_u01(seed, c)).astype(np.float32).astype(np.float64)
    v = (lo + (hi - lo) * splitmix_u01(seed, c + np.uint64(1))).astype(np.float32).astype(np.float64)
    u1 = 1.0 - splitmix_u01(seed, c + np.uint64(2))
    u2 = splitmix_u01(seed, c + np.uint64(3))
    rr = np.sqrt(-2.0 * np.log(u1))
    vals = np.stack([rr * np.cos(2 * math.pi * u2), rr * np.sin(2 * math.pi * u2)], 1)
    return np.stack([u, v], 1), vals


def config_gcf(S, beta=0.05):
    sigma = S / 6.0
    sy, sx, r, c = np.meshgrid(np.arange(8), np.arange(8), np.arange(S), np.arange(S), indexing="ij")
    dx = (c - S // 2 + 1) - sx / 8.0
    dy = (r - S // 2 + 1) - sy / 8.0
    q = dx * dx + dy * dy
    amp = np.exp(-q / (2 * sigma * sigma))
    g = np.stack([amp * np.cos(beta * q), amp * np.sin(beta * q)], -1)
    return g.reshape(-1, 2)


def pct(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * p / 100.0
    lo, hi = int(math.floor(k)), int(math.ceil(k))
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def time_variant(torch, ctx, fn, warmup, steps):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ctx.profiling(True)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [a.elapsed_time(b) for a, b in evs]
    kern = ctx.kernel_ms()
    ctx.profiling(False)
    return wall, per, kern


def host_cpu():
    """The host's CPU model and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": len(os.sched_getaffinity(0)),
            "logical_cpus": os.cpu_count()}


def omp_threads():
    """OMP_NUM_THREADS, else the usable cores up to 16 (the box's CPU share)."""
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, len(os.sched_getaffinity(0)))


def oracle_mod():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the CPU restatement of the reference gridCPU (the checker)
    return O


def cpu_baseline(uv, vals, gcf, N, S, nsample, workload):
    """The reference's own parallelism: gridCPU (grid.cu:87-139) on one core,
    in its loop order, over a bounded sample of the workload."""
    O = oracle_mod()
    t0 = time.perf_counter()
    O.grid_zz(uv[:nsample], vals[:nsample], N, gcf, S)
    dt = time.perf_counter() - t0
    return dict({"value": nsample / dt / 1e6, "unit": "Mvis/s", "cores": 1, "kind": "port",
                 "sample": f"first {nsample} visibilities of the {workload} workload through the "
                           f"single-threaded gridCPU restatement (reference loop order), {dt:.1f} s"},
                **host_cpu())


def oracle_full(uv, vals, gcf, N, S, workload, r2=None, i2=None):
    """The whole workload through the row-band OpenMP form of the gridCPU
    restatement (bit-identical to the serial loop): the error oracle of every
    variant, and the threaded CPU figure.  With r2 / i2, the constant-tap
    (COMPUTE_GCF) form.  Returns (planar grid, baseline dict)."""
    O = oracle_mod()
    want = omp_threads()
    t0 = time.perf_counter()
    if r2 is None:
        ref, used = O.grid_zz_omp(uv, vals, N, gcf, S, want)
    else:
        ref, used = O.grid_const_tap_omp(uv, vals, N, S, r2, i2, want)
    dt = time.perf_counter() - t0
    n = uv.shape[0]
    return ref, dict({"value": n / dt / 1e6, "unit": "Mvis/s", "cores": int(used), "kind": "port",
                      "sample": f"the whole {workload} workload ({n} visibilities), row-band "
                                f"OpenMP form of the gridCPU restatement on {used} threads, "
                                f"{dt:.1f} s"}, **host_cpu())


# COMPUTE_GCF reads T and w from the table element S(S-1)/2-1
# (grid_gpu.cu:103-108); the benchmark GCF holds 1 + 0i there (w = 0, a
# constant tap), so the computed variants use a copy holding T = 2, w = 0.31.
COMPAT_T, COMPAT_W = 2.0, 0.31


def compat_table(gcf, S):
    g = np.array(gcf, np.float64, copy=True)
    g[S * (S - 1) // 2 - 1] = (COMPAT_T, COMPAT_W)
    return g


def gen_clustered_dev(torch, npts, N, S, mean_r, gen):
    """Clustered uv and N(0,1) values drawn on the device (torch generator):
    angle U[0, 2pi), radius Exp(mean_r) around the centre, redrawn until the
    support window lies on the grid, uv rounded to float32 values."""
    h = S // 2
    uv = torch.empty((npts, 2), dtype=torch.float64, device="cuda")
    done = 0
    while done < npts:
        m = int((npts - done) * 1.3) + 1024
        th = 2 * math.pi * torch.rand(m, dtype=torch.float64, device="cuda", generator=gen)
        r = -mean_r * torch.log1p(-torch.rand(m, dtype=torch.float64, device="cuda", generator=gen))
        u = (N / 2 + r * torch.cos(th)).float().double()
        v = (N / 2 + r * torch.sin(th)).float().double()
        fu, fv = torch.floor(u), torch.floor(v)
        ok = (fu - h + 1 >= 0) & (fu + h <= N - 1) & (fv - h + 1 >= 0) & (fv + h <= N - 1)
        k = min(int(ok.sum()), npts - done)
        uv[done:done + k] = torch.stack([u[ok][:k], v[ok][:k]], 1)
        done += k
    vals = torch.randn((npts, 2), dtype=torch.float64, device="cuda", generator=gen)
    return uv, vals


def batch_jobs(rng, njobs):
    """Job shapes of the batched workload: visibilities log-uniform in
    [50k, 2M], grid edge from {1024, 2048, 4096}, support from {8, 16, 32, 64}."""
    n = np.exp(rng.uniform(math.log(5e4), math.log(2e6), njobs)).astype(np.int64)
    Ns = rng.choice([1024, 2048, 4096], njobs)
    Ss = rng.choice([8, 16, 32, 64], njobs)
    return [(int(a), int(b), int(c)) for a, b, c in zip(Ns, Ss, n)]


def job_roofline_ms(N, S, n, accumulate, fp32=True):
    """The roofline time of one gridding job: the larger of its VALU time
    (8 S^2 flops per visibility at the vector peak) and its HBM time
    (32 B of FP64 inputs per visibility, the 64 S^2 complex128 table, the
    N^2 complex128 grid written once, or read and written when accumulating)."""
    flops = 8.0 * S * S * n
    bytes_ = 32.0 * n + 64.0 * S * S * 16 + 16.0 * N * N * (2 if accumulate else 1)
    peak = PEAK_FP32_TF if fp32 else PEAK_FP64_TF
    return max(flops / (peak * 1e12), bytes_ / (PEAK_HBM_GBS * 1e9)) * 1e3, flops


def mix_roofline(parts, ms, accumulate):
    """roofline of a mix of jobs (N, S, n) run in ms: achieved TFLOP/s against
    the FP32 vector peak, and the fraction of the per-job roofline time
    (time-weighted: the jobs' roofline times summed, over the measured time)."""
    t_roof = flops = 0.0
    for N, S, n in parts:
        t, f = job_roofline_ms(N, S, n, accumulate)
        t_roof += t
        flops += f
    achieved = flops / (ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TF, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TF, "traffic": None,
            "per_job_roofline_ms": t_roof, "frac_of_per_job_roofline": t_roof / ms,
            "flops": flops,
            "rule": "per job max(8 S^2 n / 157.3 TFLOP/s, bytes / 8 TB/s), summed over jobs"}


def cpu_mix_baseline(parts_host, gcfs, sample, what):
    """gridCPU restatement, single thread, on the first `sample` visibilities
    of each job; the mix's rate is its visibilities over the jobs' sampled
    times scaled to their sizes."""
    O = oracle_mod()
    t_tot = 0.0
    n_tot = 0
    sampled = 0
    for (uv, vals, N, S) in parts_host:
        m = min(sample, uv.shape[0])
        if m == 0:
            continue
        t0 = time.perf_counter()
        O.grid_zz(uv[:m], vals[:m], N, gcfs[S], S)
        dt = time.perf_counter() - t0
        t_tot += dt * uv.shape[0] / m
        n_tot += uv.shape[0]
        sampled += m
    return dict({"value": n_tot / t_tot / 1e6, "unit": "Mvis/s", "cores": 1, "kind": "port",
                 "sample": f"first {sample} visibilities of each of the {len(parts_host)} {what}, "
                           f"{sampled} in all, through the single-threaded gridCPU restatement; each "
                           f"job's time scaled to its size and summed"}, **host_cpu())


def main_batch(args, torch, migrid):
    rng = np.random.default_rng(args.seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(args.seed)
    shapes = batch_jobs(rng, args.jobs)
    gcf_h = {S: config_gcf(S) for S in (8, 16, 32, 64)}
    gcfs = {S: torch.from_numpy(g).cuda() for S, g in gcf_h.items()}
    jobs = []
    for N, S, n in shapes:
        uv, vals = gen_clustered_dev(torch, n, N, S, 800.0 * N / 8192, gen)
        jobs.append(dict(grid=torch.empty((N, N, 2), dtype=torch.float64, device="cuda"),
                         uv=uv, vals=vals, gcf=gcfs[S], N=N, S=S))
    torch.cuda.synchronize()
    total = sum(n for _, _, n in shapes)
    print(f"# batch: {len(jobs)} jobs, {total} visibilities", file=sys.stderr, flush=True)
    ctx = migrid.Context()
    t0 = time.perf_counter()
    best, best_ms = migrid.autotune(ctx, jobs, arith=migrid.ARITH_F32, tap_smax=args.tap_smax)
    tune_s = time.perf_counter() - t0
    print(f"# autotune: tile {best.tile} vis/item {best.vis_per_wave} wg/CU {best.wg_per_cu} "
          f"stage_min {best.stage_min} {best_ms:.2f} ms ({tune_s:.1f} s)", file=sys.stderr, flush=True)
    b = migrid.Batch(ctx, jobs, opts=best)
    stream = torch.cuda.Stream()  # graph capture needs a non-default stream
    for _ in range(args.warmup):
        b.graph_launch(stream)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        b.graph_launch(stream)
        e1.record(stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    per = [e0.elapsed_time(e1) for e0, e1 in evs]
    # the checked result: one more replay into NaN-filled grids (untimed)
    for j in jobs:
        j["grid"].fill_(float("nan"))
    torch.cuda.synchronize()
    b.graph_launch(stream)
    torch.cuda.synchronize()
    host = [(j["uv"].cpu().numpy(), j["vals"].cpu().numpy(), j["N"], j["S"]) for j in jobs]
    err = {"rel_max_vs_oracle": 0.0, "rel_rms_vs_oracle": 0.0}
    base_thr = None
    if not args.no_cpu:
        # every job through the OpenMP gridCPU restatement: the error oracle,
        # and the threaded CPU figure for the whole batch
        O = oracle_mod()
        t_cpu, used = 0.0, 0
        for j, (uv, vals, N, S) in zip(jobs, host):
            t1 = time.perf_counter()
            ref, used = O.grid_zz_omp(uv, vals, N, gcf_h[S], S, omp_threads())
            t_cpu += time.perf_counter() - t1
            r = torch.from_numpy(np.ascontiguousarray(np.moveaxis(ref, 0, -1))).cuda()
            d = torch.linalg.vector_norm(j["grid"] - r, dim=-1)
            m = torch.linalg.vector_norm(r, dim=-1)
            err["rel_max_vs_oracle"] = max(err["rel_max_vs_oracle"], float(d.max() / m.max()))
            err["rel_rms_vs_oracle"] = max(err["rel_rms_vs_oracle"],
                                           float(torch.sqrt((d * d).sum() / (m * m).sum())))
            del r, d, m, ref
        base_thr = dict({"value": total / t_cpu / 1e6, "unit": "Mvis/s", "cores": int(used),
                         "kind": "port",
                         "sample": f"every job of the batch ({total} visibilities), row-band "
                                   f"OpenMP form of the gridCPU restatement on {used} threads, "
                                   f"{t_cpu:.1f} s"}, **host_cpu())
        print(f"# batch oracle: {base_thr['sample']}; worst job {err}", file=sys.stderr, flush=True)
    # the plain FP64 path: one scatter launch per job
    t0 = time.perf_counter()
    for j in jobs:
        ctx.grid(j["grid"], j["uv"], j["vals"], j["gcf"], j["N"], j["S"], strategy="scatter")
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3
    out = {
        "metric": "Mvis/s, 64 uneven gridding jobs in one hipGraph-captured persistent launch",
        "value": total / (ms * 1e-3) / 1e6, "unit": "Mvis/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(jobs)} jobs: vis log-uniform [50k, 2M], grids "
                               "{1024,2048,4096}^2, supports {8,16,32,64}, clustered uv, "
                               "FP32 tiles into FP64 grids", "seed": args.seed, "npts_total": total},
        "p50_ms": pct(per, 50), "p99_ms": pct(per, 99),
        "autotune": {"tile": best.tile, "vis_per_item": best.vis_per_wave,
                     "wg_per_cu": best.wg_per_cu, "stage_min": best.stage_min, "ms": best_ms,
                     "seconds": tune_s},
        "fp64_plain_one_launch_per_job": {"ms": plain_ms,
                                          "mvis_per_s": total / (plain_ms * 1e-3) / 1e6},
        "err_worst_job": err if not args.no_cpu else None,
        "roofline": mix_roofline(shapes, ms, accumulate=False),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_mix_baseline(host, gcf_h, args.cpu_sample // 25,
                                               "jobs of the batch")
        out["cpu_baseline_threads"] = base_thr
    print(json.dumps(out), flush=True)


def main_serve(args, torch, migrid):
    import threading
    rng = np.random.default_rng(args.seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(args.seed)
    gcf_h = {S: config_gcf(S) for S in (8, 16, 32, 64)}
    gcfs = {S: torch.from_numpy(g).cuda() for S, g in gcf_h.items()}
    small = batch_jobs(rng, args.jobs)  # (N, S, max vis) per small job id
    pools = {}
    for jid, (N, S, _) in enumerate(small):
        pools[jid] = gen_clustered_dev(torch, 2_000_000, N, S, 800.0 * N / 8192, gen) + (N, S)
    nlarge = args.large_jobs
    big = gen_clustered_dev(torch, 10_000_000, 8192, 64, 800.0, gen)
    for k in range(nlarge):
        pools[1000 + k] = big + (8192, 64)
    torch.cuda.synchronize()
    # Poisson arrivals; 10% of requests are 8192^2 / 10M visibilities
    n_req = int(args.rate * args.duration * 1.2) + 16
    arr = np.cumsum(rng.exponential(1.0 / args.rate, n_req))
    arr = arr[arr < args.duration]
    trace = []
    for t in arr:
        if rng.uniform() < 0.1:
            jid, n, off = 1000 + int(rng.integers(nlarge)), 10_000_000, 0
        else:
            jid = int(rng.integers(len(small)))
            n = int(math.exp(rng.uniform(math.log(5e4), math.log(2e6))))
            off = int(rng.integers(0, 2_000_000 - n + 1))
        trace.append((float(t), jid, n, off))
    total = sum(n for _, _, n, _ in trace)
    print(f"# serve: {len(trace)} requests over {args.duration} s, {total} visibilities",
          file=sys.stderr, flush=True)

    def req_args(jid, n, off):
        uv, vals, N, S = pools[jid]
        return uv[off:off + n], vals[off:off + n], gcfs[S], N, S

    srv = migrid.Server(max_batch=args.max_batch, max_inflight=args.max_inflight)
    # warm-up request on a job id outside the trace
    srv.wait(srv.submit(-1, *req_args(0, 50_000, 0)))
    srv.drop_job(-1)
    st0 = srv.stats()
    tickets = []
    kept = set()
    t_start = time.perf_counter()
    for t, jid, n, off in trace:
        now = time.perf_counter() - t_start
        if t > now:
            time.sleep(t - now)
        # the large job ids repeat their coordinates: the first request keeps
        # its sort, later ones reuse it
        flags = {}
        if jid >= 1000:
            flags = {"same_uv": True} if jid in kept else {"keep_sort": True}
            kept.add(jid)
        tickets.append(srv.submit(jid, *req_args(jid, n, off), **flags))
        if len(tickets) % 2000 == 0:
            print(f"# serve: {len(tickets)} submitted", file=sys.stderr, flush=True)
    lat = [srv.wait(tk) for tk in tickets]
    wall = time.perf_counter() - t_start
    st = srv.stats()
    srv.close()
    batches = st["batches"] - st0["batches"]
    busy_ms = st["gpu_busy_ms"] - st0["gpu_busy_ms"]
    # the plain FP64 path on a prefix of the same trace: FIFO, one scatter
    # launch per request, accumulating into per-job grids
    prefix = [r for r in trace if r[0] < args.plain_duration]
    ctx = migrid.Context()
    grids = {}
    plain_lat = []
    t_start = time.perf_counter()
    for t, jid, n, off in prefix:
        now = time.perf_counter() - t_start
        if t > now:
            time.sleep(t - now)
        uv, vals, g, N, S = req_args(jid, n, off)
        if jid not in grids:
            grids[jid] = torch.zeros((N, N, 2), dtype=torch.float64, device="cuda")
        ctx.grid(grids[jid], uv, vals, g, N, S, strategy="scatter", accumulate=True)
        torch.cuda.synchronize()
        plain_lat.append((time.perf_counter() - t_start - t) * 1e3)
    plain_wall = time.perf_counter() - t_start
    plain_vis = sum(n for _, _, n, _ in prefix)
    parts = [(pools[jid][2], pools[jid][3], n) for _, jid, n, _ in trace]
    out = {
        "metric": f"Mvis/s and p99 latency of the serving loop, Poisson {args.rate:g} req/s",
        "value": total / wall / 1e6, "unit": "Mvis/s", "n_gpus": 1, "steps": len(trace),
        "warmup": 1, "ms_per_step": wall * 1e3 / len(trace), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"Poisson {args.rate} req/s for {args.duration} s; 90% of requests "
                               f"on {len(small)} job ids with grids {{1024,2048,4096}}^2, supports "
                               "{8,16,32,64}, vis log-uniform [50k, 2M]; 10% on "
                               f"{nlarge} job ids of 8192^2, S=64, 10M vis; clustered uv; FP32 "
                               "tiles into resident FP64 grids", "seed": args.seed,
                   "requests": len(trace), "npts_total": total},
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "max_ms": max(lat),
        "batches": batches, "mean_requests_per_batch": len(trace) / max(batches, 1),
        "gpu_busy_frac": busy_ms / (wall * 1e3),
        "gpu_overlap_ms": st["overlap_ms"] - st0["overlap_ms"],
        "batch_device_ms_summed": st["gpu_ms"] - st0["gpu_ms"],
        "sorts_reused": st["sorts_reused"] - st0["sorts_reused"],
        "offered_mvis_per_s": total / args.duration / 1e6,
        # against the device-busy time: how close the served work runs to its
        # per-request roofline while the GPU is working
        "roofline": mix_roofline(parts, busy_ms, accumulate=True),
        "fp64_plain_one_launch_per_request": {
            "requests": len(prefix), "trace_seconds": args.plain_duration,
            "mvis_per_s": plain_vis / plain_wall / 1e6,
            "p50_ms": pct(plain_lat, 50), "p99_ms": pct(plain_lat, 99)},
    }
    if not args.no_cpu:
        # gridCPU restatement, single thread, on the first visibilities of each
        # job id, each job id's time scaled to the visibilities the trace sends it
        O = oracle_mod()
        load = {}
        for _, jid, n, _ in trace:
            load[jid] = load.get(jid, 0) + n
        m = args.cpu_sample // 25
        t_tot = 0.0
        for jid in sorted(load):
            uv, vals, N, S = pools[jid]
            uv_h, vals_h = uv[:m].cpu().numpy(), vals[:m].cpu().numpy()
            t1 = time.perf_counter()
            O.grid_zz(uv_h, vals_h, N, gcf_h[S], S)
            t_tot += (time.perf_counter() - t1) * load[jid] / m
        cb = dict({"value": total / t_tot / 1e6, "unit": "Mvis/s", "cores": 1, "kind": "port",
                   "sample": f"first {m} visibilities of each of the {len(load)} job ids through "
                             "the single-threaded gridCPU restatement, each job id's time scaled "
                             "to the visibilities the trace sends it"}, **host_cpu())
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


def main_degrid(args, torch, migrid):
    """Degridding (the adjoint of gridding) of the gridding workload: the
    same uv onto a complex N(0,1) grid drawn on the device, FP64 and FP32
    products.  A step is one full pass: zero fill, sort by (tile, sub-kernel)
    and the degridding kernel."""
    N, S, npts = args.N, args.S, args.npts
    uv, _ = gen_clustered(npts, N, S) if args.dist == "clustered" else gen_uniform(npts, N, S)
    gcf = config_gcf(S)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(args.seed)
    d_uv = torch.from_numpy(uv).cuda()
    d_gcf = torch.from_numpy(gcf).cuda()
    grid = torch.randn((N, N, 2), dtype=torch.float64, device="cuda", generator=gen)
    out = torch.empty((npts, 2), dtype=torch.float64, device="cuda")
    ctx = migrid.Context()
    nchk = min(npts, 12000)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # CPU restatement, the checker of the first visibilities
    hgrid = grid.permute(2, 0, 1).contiguous().cpu().numpy()
    t_cpu = time.perf_counter()
    ref = O.degrid_zz(uv[:nchk], hgrid, N, gcf, S)
    t_cpu = time.perf_counter() - t_cpu
    results = {}
    for name, arith in (("fp64", migrid.ARITH_NATIVE), ("fp32", migrid.ARITH_F32)):
        def step(arith=arith):
            ctx.degrid(grid, d_uv, d_gcf, N, S, out=out, arith=arith, tile=args.tile,
                       vis_per_wave=args.vpw, wg_per_cu=args.wg_per_cu)
        try:
            step()
        except migrid.MigridError as e:
            print(f"# degrid {name}: not available ({e})", file=sys.stderr, flush=True)
            continue
        print(f"# degrid {name}: timing {args.steps} steps", file=sys.stderr, flush=True)
        wall, per, kern = time_variant(torch, ctx, step, args.warmup, args.steps)
        got = out[:nchk].cpu().numpy()
        err = float(np.abs(got - ref).max() / np.abs(ref).max())
        ms = wall * 1e3 / args.steps
        results[name] = {"ms_per_step": ms, "mvis_per_s": npts / (ms * 1e-3) / 1e6,
                         "p50_ms": pct(per, 50), "p99_ms": pct(per, 99),
                         "kernel_ms_avg": float(np.mean(kern)) if kern else None,
                         "rel_max_vs_oracle_first_%d" % nchk: err}
        print(f"# degrid {name}: {results[name]}", file=sys.stderr, flush=True)
    head = "fp32" if "fp32" in results else "fp64"
    R = results[head]
    flops = 8.0 * S * S * npts
    achieved = flops / (R["kernel_ms_avg"] * 1e-3) / 1e12
    out_line = {
        "metric": f"Mvis/s degridding at {N}² grid, support {S}",
        "value": R["mvis_per_s"], "unit": "Mvis/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": R["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"degridding {N}x{N} complex128 grid, support {S}, oversampling 8, "
                               f"{npts} radially clustered visibilities (Exp radius mean 800 px)",
                   "variant": head},
        "variants": results,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TF,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TF, "traffic": None,
                     "kernel": "degridding kernel", "flops_per_launch": flops},
        "cpu_baseline": dict({"value": nchk / t_cpu / 1e6, "unit": "Mvis/s", "cores": 1,
                              "kind": "port",
                              "sample": f"first {nchk} visibilities of the same workload through "
                                        f"the oracle's single-threaded degridder (the adjoint of "
                                        f"the gridCPU restatement), {t_cpu:.1f} s"}, **host_cpu()),
    }
    print(json.dumps(out_line), flush=True)


def main_image(args, torch, migrid):
    """Grid -> image: the centred 2-D FFT of an N x N complex grid (the step
    after gridding).  A step is one full transform: sign/layout pass, hipFFT
    and the sign/scale/taper pass.  Algorithmic bytes: the grid read once and
    the image written once."""
    N = args.N
    gen = torch.Generator(device="cuda")
    gen.manual_seed(args.seed)
    ctx = migrid.Context()
    taper = torch.linspace(0.5, 1.0, N, dtype=torch.float64, device="cuda")
    results = {}
    for name, dt, esz in (("fp64", torch.float64, 16), ("fp32", torch.float32, 8)):
        g = torch.randn((N, N, 2), dtype=dt, device="cuda", generator=gen)
        out = torch.empty_like(g)

        def step(g=g, out=out):
            ctx.fft(g, out=out, scale=1.0 / (N * N), taper=taper)
        print(f"# image {name}: timing {args.steps} steps", file=sys.stderr, flush=True)
        wall, per, kern = time_variant(torch, ctx, step, max(args.warmup, 1), args.steps)
        ms = wall * 1e3 / args.steps
        kms = float(np.mean(kern))
        results[name] = {"ms_per_step": ms, "images_per_s": 1e3 / ms, "p50_ms": pct(per, 50),
                         "p99_ms": pct(per, 99), "transform_ms_avg": kms,
                         "algorithmic_gbs": 2 * N * N * esz / (kms * 1e-3) / 1e9}
        print(f"# image {name}: {results[name]}", file=sys.stderr, flush=True)
        del g, out
    R = results["fp64"]
    out_line = {
        "metric": f"grid to image transforms/s at {N}², complex128",
        "value": R["images_per_s"], "unit": "images/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": R["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"centred 2-D FFT of a {N}x{N} complex128 grid with taper "
                               f"correction and scale", "variant": "fp64"},
        "variants": results,
        "roofline": {"bound": "hbm", "achieved": R["algorithmic_gbs"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": R["algorithmic_gbs"] / PEAK_HBM_GBS, "traffic": None,
                     "kernel": "pre pass + hipFFT + post pass",
                     "bytes_per_launch": 2 * N * N * 16},
    }
    if not args.no_cpu:
        # the oracle's numpy restatement of the centred transform, same grid
        O = oracle_mod()
        g = torch.randn((N, N, 2), dtype=torch.float64, device="cuda", generator=gen)
        gh = g[..., 0].cpu().numpy() + 1j * g[..., 1].cpu().numpy()
        del g
        t_cpu = time.perf_counter()
        O.grid_to_image(gh, 1.0 / (N * N), np.linspace(0.5, 1.0, N))
        t_cpu = time.perf_counter() - t_cpu
        out_line["cpu_baseline"] = dict({"value": 1.0 / t_cpu, "unit": "images/s", "cores": 1,
                                         "kind": "port",
                                         "sample": f"one {N}x{N} complex128 transform with taper "
                                                   f"through numpy's FFT (the oracle's "
                                                   f"restatement), {t_cpu:.1f} s"}, **host_cpu())
    print(json.dumps(out_line), flush=True)


def main_ceilings(args, torch, migrid):
    """Measured ceilings beside the kernels they bound, on the grid workload:
    LDS read bandwidth (the tiled owner kernel reads one GCF element from LDS
    per executed tap) and global float atomics at the plain scatter kernel's
    address distribution (2 S^2 per visibility)."""
    N, S, npts = args.N, args.S, args.npts
    uv, vals = gen_clustered(npts, N, S) if args.dist == "clustered" else gen_uniform(npts, N, S)
    gcf = config_gcf(S)
    d_uv = torch.from_numpy(uv).cuda()
    d_vals = torch.from_numpy(vals).cuda()
    d_gcf = torch.from_numpy(gcf).cuda()
    lds = {}
    for eb in (8, 16):
        for pitch in (0, 96, 100, 104, 108, 112, 116, 120, 124, 97, 99):
            r = migrid.ceiling_lds(eb, pitch, iters=args.lds_iters)
            r["tb_per_s"] = r["per_s"] / 1e12
            pname = "linear" if pitch == 0 else f"owner_block_pitch{pitch}"
            lds[f"b{8 * eb}_{pname}"] = r
            print(f"# lds {eb} B {pname}: {r['tb_per_s']:.1f} TB/s ({r['ms']:.3f} ms)",
                  file=sys.stderr, flush=True)
    atom = {}
    ctx = migrid.Context()
    kern = {}
    for name, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        g = torch.zeros((N, N, 2), dtype=dt, device="cuda")
        migrid.ceiling_atomics(g, d_uv, N, S)  # warm-up
        r = migrid.ceiling_atomics(g, d_uv, N, S)
        r["atomics_per_vis"] = 2 * S * S
        atom[name] = r
        print(f"# atomics {name}: {r}", file=sys.stderr, flush=True)
        # the plain scatter kernel on the same inputs and grid type
        uv_in = d_uv if dt == torch.float64 else d_uv.float()
        vals_in = d_vals if dt == torch.float64 else d_vals.float()
        gcf_in = d_gcf if dt == torch.float64 else d_gcf.float()
        _, _, km = time_variant(torch, ctx, lambda: ctx.grid(g, uv_in, vals_in, gcf_in, N, S,
                                                             strategy="scatter"), 1, 2)
        kern[f"scatter_{name}"] = float(np.mean(km))
        del g
    # the owner kernels on the same workload: executed taps from the bin
    # geometry (each visibility listed by (ceil((S + 15) / 16))^2 waves of 256
    # lane-taps when its window is inside the grid), one LDS element per
    # lane-tap: four 16-byte reads per visibility and wave for FP64 pairs
    # (ds_read_b128) and two for FP32 pairs (ds_read2_b64, two 8-byte
    # elements each).  The probe shows LDS time goes per wave instruction
    # (8- and 16-byte reads take the same time when free of bank conflicts),
    # so the ceiling is the 16-byte probe's instruction rate at the kernel's
    # pitch (104 elements, 8 mod 32).
    waves = ((S + 15 + 15) // 16) ** 2
    instr_ceiling = lds["b128_owner_block_pitch104"]["per_s"] / (64 * 16)
    g = torch.empty((N, N, 2), dtype=torch.float64, device="cuda")
    owner = {}
    for name, arith, eb, per_vis in (("fp64", migrid.ARITH_NATIVE, 16, 4),
                                     ("fp32", migrid.ARITH_F32, 8, 2)):
        _, _, km = time_variant(torch, ctx, lambda: ctx.grid(g, d_uv, d_vals, d_gcf, N, S,
                                                             strategy="tiled", arith=arith), 2, 5)
        ms = float(np.mean(km))
        lds_bytes = float(npts) * waves * 256 * eb
        instr = float(npts) * waves * per_vis
        owner[name] = {"kernel_ms": ms, "lds_bytes_est": lds_bytes,
                       "lds_tb_per_s": lds_bytes / (ms * 1e-3) / 1e12,
                       "lds_wave_instr_est": instr,
                       "frac_of_lds_instr_ceiling": instr / (ms * 1e-3) / instr_ceiling,
                       "executed_taps_per_vis": waves * 256, "useful_taps_per_vis": S * S}
    out = {"metric": "measured ceilings: LDS read bandwidth and global float atomics",
           "config": {"workload": f"{N}x{N} grid, support {S}, {npts} {args.dist} visibilities"},
           "lds": lds, "lds_wave_instr_per_s_ceiling": instr_ceiling, "atomics": atom,
           "owner_kernel": owner,
           "scatter_kernel": {k: {"kernel_ms": v,
                                  "frac_of_atomic_ceiling": atom[k.split("_")[1]]["ms"] / v}
                              for k, v in kern.items()}}
    print(json.dumps(out), flush=True)


def main_copy(args, torch):
    """Plain device-to-device copy of 1 GiB: the calibration kernel for the
    HBM byte counters (it reads and writes exactly 2^30 bytes per launch)."""
    n = 2**28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    print(json.dumps({"metric": "device copy GB/s (read + write)", "value": 2 * 4 * n / ms / 1e6,
                      "unit": "GB/s", "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
                      "ms_per_launch": ms, "steps": args.steps}), flush=True)


def load_traffic():
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--npts", type=int, default=10_000_000)
    ap.add_argument("--N", type=int, default=8192)
    ap.add_argument("--S", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=100000,
                    help="visibilities of the single-threaded CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variants", default="fp64_plain,fp64_tiled,fp32_tiled,fp32_tiled_twosum,"
                                          "bf16_tiled,fp64_gather,fp64_window,fp64_plain_computed,"
                                          "fp64_tiled_computed")
    ap.add_argument("--tile", type=int, default=0, help="tiled: tile edge 16/32/64 (0 auto)")
    ap.add_argument("--vpw", type=int, default=0, help="tiled: visibilities per work item")
    ap.add_argument("--stage-min", type=int, default=0, help="tiled: staging threshold")
    ap.add_argument("--wg-per-cu", type=int, default=0, help="tiled: workgroups per CU")
    ap.add_argument("--tap-smax", type=int, default=0,
                    help="tiled: largest support gridded one tap per lane (0: 16, -1: never)")
    ap.add_argument("--dist", default="clustered", choices=["clustered", "uniform"])
    ap.add_argument("--mode", default="grid", choices=["grid", "batch", "serve", "copy", "degrid", "image", "cpu",
                             "ceilings"],
                    help="grid: one 8192^2 job (default); batch: many jobs in one persistent "
                         "launch; serve: Poisson request stream through the serving loop; "
                         "copy: 1 GiB device copy (byte-counter calibration); degrid: the "
                         "adjoint (grid -> visibilities) of the grid workload; image: the "
                         "grid -> image FFT; cpu: the CPU baselines in full (no GPU); "
                         "ceilings: LDS read and global atomic ceilings beside the kernels")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--jobs", type=int, default=64, help="batch/serve: job ids")
    ap.add_argument("--large-jobs", type=int, default=4, help="serve: 8192^2 job ids")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: requests per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--plain-duration", type=float, default=1.0,
                    help="serve: seconds of the trace replayed on the plain FP64 path")
    ap.add_argument("--max-batch", type=int, default=0)
    ap.add_argument("--max-inflight", type=int, default=0)
    ap.add_argument("--lds-iters", type=int, default=20000, help="ceilings: LDS probe steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus %d is not supported "
                                   "(multi-GPU sharding is described, not built)" % args.gpus}))
        sys.exit(2)

    if args.mode == "cpu":
        return main_cpu(args)
    import torch
    import migrid

    if args.mode == "batch":
        return main_batch(args, torch, migrid)
    if args.mode == "serve":
        return main_serve(args, torch, migrid)
    if args.mode == "copy":
        return main_copy(args, torch)
    if args.mode == "degrid":
        return main_degrid(args, torch, migrid)
    if args.mode == "image":
        return main_image(args, torch, migrid)
    if args.mode == "ceilings":
        return main_ceilings(args, torch, migrid)
    N, S, npts = args.N, args.S, args.npts
    t0 = time.perf_counter()
    uv, vals = gen_clustered(npts, N, S) if args.dist == "clustered" else gen_uniform(npts, N, S)
    gcf = config_gcf(S)
    gcf_c = compat_table(gcf, S)
    gen_s = time.perf_counter() - t0
    print(f"# generated {npts} visibilities in {gen_s:.1f} s", file=sys.stderr, flush=True)
    workload = f"{N}^2/S={S}/{npts} {args.dist}"

    d_uv = torch.from_numpy(uv).cuda()
    d_vals = torch.from_numpy(vals).cuda()
    d_gcf = torch.from_numpy(gcf).cuda()
    d_gcf_c = torch.from_numpy(gcf_c).cuda()
    grid = torch.empty((N, N, 2), dtype=torch.float64, device="cuda")
    grid32 = grid_lo = None
    ctx = migrid.Context()

    variants = {
        "fp64_plain": dict(strategy="scatter"),
        "fp64_tiled": dict(strategy="tiled", arith=migrid.ARITH_NATIVE),
        "fp32_tiled": dict(strategy="tiled", arith=migrid.ARITH_F32),
        # FP32 tiles flushed into an FP32 grid plus its two-sum compensation grid
        "fp32_tiled_twosum": dict(strategy="tiled", arith=migrid.ARITH_F32, out32=True),
        "bf16_tiled": dict(strategy="tiled", arith=migrid.ARITH_F32_BF16GCF),
        # the reference's other two strategies, FP64 (interleaved for the comparison)
        "fp64_gather": dict(strategy="gather"),
        "fp64_window": dict(strategy="window", layout=migrid.LAYOUT_INTERLEAVED),
        # COMPUTE_GCF: the phase form evaluated per visibility, scatter and tiled
        "fp64_plain_computed": dict(strategy="scatter", gcf_mode=migrid.GCF_COMPAT),
        "fp64_tiled_computed": dict(strategy="tiled", gcf_mode=migrid.GCF_COMPAT),
    }
    tune = dict(tile=args.tile, vis_per_wave=args.vpw, stage_min=args.stage_min,
                wg_per_cu=args.wg_per_cu, tap_smax=args.tap_smax)
    names = args.variants.split(",")
    # the FP64 oracle of the whole workload, on the host's cores (also the
    # threaded CPU figure); the constant-tap oracle for the COMPUTE_GCF variants
    oracle = {}
    if not args.no_cpu:
        print(f"# oracle: gridCPU restatement over the whole workload on {omp_threads()} threads",
              file=sys.stderr, flush=True)
        ref, base_thr = oracle_full(uv, vals, gcf, N, S, workload)
        oracle["table"] = torch.from_numpy(np.ascontiguousarray(np.moveaxis(ref, 0, -1))).cuda()
        del ref
        print(f"# oracle: {base_thr['sample']}", file=sys.stderr, flush=True)
        if any(variants[n].get("gcf_mode") == migrid.GCF_COMPAT for n in names):
            O = oracle_mod()
            r2, i2 = O.compute_gcf_phase(uv, COMPAT_T, COMPAT_W, fixed_point=False)
            ref, _ = oracle_full(uv, vals, gcf, N, S, workload, r2, i2)
            oracle["compat"] = torch.from_numpy(np.ascontiguousarray(np.moveaxis(ref, 0, -1))).cuda()
            del ref, r2, i2
    results = {}
    for name in names:
        kw = dict(variants[name])
        out32 = kw.pop("out32", False)
        if kw["strategy"] != "scatter":
            kw.update(tune)
        computed = kw.get("gcf_mode", migrid.GCF_TABLE) == migrid.GCF_COMPAT
        g_in = d_gcf_c if computed else d_gcf
        # the plain scatter runs ~1 s a call: fewer steps (the headline keeps K, W)
        plain = kw["strategy"] == "scatter"
        warmup, steps = (min(args.warmup, 1), min(args.steps, 3)) if plain else (args.warmup, args.steps)
        if out32 and grid32 is None:
            grid32 = torch.empty((N, N, 2), dtype=torch.float32, device="cuda")
            grid_lo = torch.empty_like(grid32)

        def step(kw=kw, g_in=g_in, out32=out32):
            if out32:
                ctx.grid(grid32, d_uv, d_vals, g_in, N, S, grid_lo=grid_lo, **kw)
            else:
                ctx.grid(grid, d_uv, d_vals, g_in, N, S, **kw)
        # the checked result: one untimed call into NaN-filled outputs, so a
        # tile the kernels leave unwritten cannot keep an earlier answer
        (grid32 if out32 else grid).fill_(float("nan"))
        if out32:
            grid_lo.fill_(float("nan"))
        try:
            step()
        except migrid.MigridError as e:
            print(f"# {name}: not available ({e})", file=sys.stderr, flush=True)
            continue
        torch.cuda.synchronize()
        err = None
        ref = oracle.get("compat" if computed else "table")
        if ref is not None:
            g = grid32.double() + grid_lo.double() if out32 else grid
            d = torch.linalg.vector_norm(g - ref, dim=-1)
            m = torch.linalg.vector_norm(ref, dim=-1)
            err = {"rel_max_vs_oracle": float(d.max() / m.max()),
                   "rel_rms_vs_oracle": float(torch.sqrt((d * d).sum() / (m * m).sum()))}
            del g, d, m
        print(f"# {name}: timing {steps} steps", file=sys.stderr, flush=True)
        wall, per, kern = time_variant(torch, ctx, step, warmup, steps)
        ms = wall * 1e3 / steps
        results[name] = {"ms_per_step": ms, "mvis_per_s": npts / (ms * 1e-3) / 1e6,
                         "grid_point_updates_per_s": npts * S * S / (ms * 1e-3),
                         "p50_ms": pct(per, 50), "p99_ms": pct(per, 99),
                         "kernel_ms_avg": float(np.mean(kern)) if kern else None,
                         "kernel_launches": len(kern), "steps": steps, "err": err}
        print(f"# {name}: {results[name]}", file=sys.stderr, flush=True)
    if not results:
        print(json.dumps({"error": "no variant ran"}))
        sys.exit(1)

    head = "fp32_tiled" if "fp32_tiled" in results else next(iter(results))
    R = results[head]
    flops = 8.0 * S * S * npts
    kms = R["kernel_ms_avg"]
    fp32 = head in ("fp32_tiled", "fp32_tiled_twosum", "bf16_tiled")
    peak = PEAK_FP32_TF if fp32 else PEAK_FP64_TF
    achieved = flops / (kms * 1e-3) / 1e12 if kms else None
    traffic = load_traffic().get(head)
    out = {
        "metric": METRIC,
        "value": R["mvis_per_s"],
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if fp32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"{N}x{N} complex128 grid, support {S}, oversampling 8, "
                               f"{npts} radially clustered visibilities (Exp radius mean 800 px)"
                               if args.dist == "clustered" else
                               f"{N}x{N} complex128 grid, support {S}, oversampling 8, "
                               f"{npts} uniform visibilities",
                   "variant": head, "grid": N, "support": S, "npts": npts},
        "p50_ms": R["p50_ms"],
        "p99_ms": R["p99_ms"],
        "grid_point_updates_per_s": R["mvis_per_s"] * 1e6 * S * S,
        "variants": results,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak if achieved else None, "traffic": traffic,
                     "kernel": "tiled owner kernel" if head != "fp64_plain" else "plain scatter kernel",
                     "flops_per_launch": flops},
    }
    if "fp64_plain" in results and head != "fp64_plain":
        out["speedup_vs_fp64_plain"] = R["mvis_per_s"] / results["fp64_plain"]["mvis_per_s"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(uv, vals, gcf, N, S, args.cpu_sample, workload)
        out["cpu_baseline_threads"] = base_thr
    print(json.dumps(out), flush=True)


def main_cpu(args):
    """The CPU baselines in full: the reference's CPU-runnable case (2048^2,
    S = 32, 1M uniform visibilities) through the single-threaded gridCPU
    restatement and its OpenMP form, and the 8192^2 / S = 64 / 10M clustered
    workload through the OpenMP form (single-threaded it would take most of
    an hour)."""
    O = oracle_mod()
    lines = {}
    uv, vals = O.gen_uniform(1_000_000, 64, 1984, seed=1541617)  # u, v in [64, 1984)
    gcf = config_gcf(32)
    t0 = time.perf_counter()
    g1 = O.grid_zz(uv, vals, 2048, gcf, 32)
    dt = time.perf_counter() - t0
    lines["2048_s32_1m_single"] = {"value": 1.0 / dt, "unit": "Mvis/s", "cores": 1, "kind": "port",
                                   "seconds": dt}
    print(f"# cpu: 2048^2/S=32/1M single thread {dt:.1f} s", file=sys.stderr, flush=True)
    gN, base = oracle_full(uv, vals, gcf, 2048, 32, "2048^2/S=32/1M uniform")
    assert np.array_equal(g1, gN)
    lines["2048_s32_1m_threads"] = base
    del g1, gN
    uv, vals = gen_clustered(args.npts, args.N, args.S)
    _, base = oracle_full(uv, vals, config_gcf(args.S), args.N, args.S,
                          f"{args.N}^2/S={args.S}/{args.npts} clustered")
    lines[f"{args.N}_s{args.S}_{args.npts}_threads"] = base
    print(json.dumps({"metric": "CPU gridCPU restatement, Mvis/s", "results": lines, **host_cpu()}),
          flush=True)


if __name__ == "__main__":
    main()
