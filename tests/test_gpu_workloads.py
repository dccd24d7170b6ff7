"""The batched and serving workloads at their own sizes, against the oracle.

* The 64-job batch: visibilities log-uniform in [50k, 2M] per job (28.6M in
  all for the seed below), grids from {1024, 2048, 4096}^2, supports from
  {8, 16, 32, 64}, radially clustered uv scaled to each grid, FP32 tiles into
  FP64 grids, planned once and launched as one persistent kernel (with the
  small supports' tap kernel): a direct run and two graph replays with eager
  work between them, every job's grid against the oracle's gridCPU within
  1e-5 relative max-norm after each, outputs filled with NaN before each.
* The serving loop on a 5-second Poisson trace at 200 requests/s: 90% of the
  requests on 64 job ids with the batch's shapes, 10% on four job ids of
  8192^2, S = 64, 10M visibilities; grids accumulate per job id.  Checked:
  every request completes; the four large grids equal their request count
  times the oracle's grid of the 10M visibilities; three small job ids'
  accumulated grids against the oracle of all their requests' visibilities;
  and one request served alone equals the same request served in the mix.
"""
import os

import numpy as np
import pytest

import bench
import migrid
import oracle as O

pytestmark = pytest.mark.gpu

NAN = float("nan")


def torch():
    import torch as t
    return t


def threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def rel_max_dev(got, ref_planar):
    """rel_max of an interleaved device grid against a planar host grid."""
    t = torch()
    ref = t.from_numpy(np.ascontiguousarray(np.moveaxis(ref_planar, 0, -1))).cuda()
    d = t.linalg.vector_norm(got.double() - ref, dim=-1)
    return float(d.max() / t.linalg.vector_norm(ref, dim=-1).max())


def rel_max_np(got, ref_planar):
    return O.rel_errors(got, ref_planar[0] + 1j * ref_planar[1])[0]


def test_64_job_batch_against_the_oracle():
    t = torch()
    rng = np.random.default_rng(2024)
    shapes = bench.batch_jobs(rng, 64)
    total = sum(n for _, _, n in shapes)
    print(f"[batch] {len(shapes)} jobs, {total} visibilities", flush=True)
    gcfs = {S: O.config_gcf(S) for S in (8, 16, 32, 64)}
    d_gcf = {S: t.from_numpy(g).cuda() for S, g in gcfs.items()}
    jobs, host = [], []
    for i, (N, S, n) in enumerate(shapes):
        uv, vals = O.gen_clustered(n, N, S, mean_r=800.0 * N / 8192, seed=1000 + i)
        host.append((uv, vals))
        jobs.append(dict(grid=t.empty((N, N, 2), dtype=t.float64, device="cuda"),
                         uv=t.from_numpy(uv).cuda(), vals=t.from_numpy(vals).cuda(),
                         gcf=d_gcf[S], N=N, S=S))
    print(f"[batch] oracle on {threads()} threads", flush=True)
    refs = [O.grid_zz_omp(uv, vals, N, gcfs[S], S, threads())[0]
            for (uv, vals), (N, S, _) in zip(host, shapes)]
    ctx = migrid.Context()
    b = migrid.Batch(ctx, jobs, arith=migrid.ARITH_F32)
    stream = t.cuda.Stream()
    extra = jobs[0]
    scratch = t.empty((extra["N"], extra["N"], 2), dtype=t.float64, device="cuda")
    worst = 0.0
    for mode in ("run", "graph", "graph"):
        for j in jobs:
            j["grid"].fill_(NAN)
        t.cuda.synchronize()
        b.run(stream) if mode == "run" else b.graph_launch(stream)
        stream.synchronize()
        for k, (j, r) in enumerate(zip(jobs, refs)):
            e = rel_max_dev(j["grid"], r)
            worst = max(worst, e)
            assert e < 1e-5, (mode, k, j["N"], j["S"], e)
        # eager work on the same context between replays
        ctx.grid(scratch, extra["uv"], extra["vals"], extra["gcf"], extra["N"], extra["S"],
                 strategy="tiled", stream=stream)
        x = t.randn(1000, device="cuda").sum()
        stream.synchronize()
        assert t.isfinite(x)
        print(f"[batch] {mode}: every job within {worst:.2e} of the oracle", flush=True)
    b.close()


def test_poisson_serving_trace_against_the_oracle():
    t = torch()
    rate, duration, nsmall, nlarge = 200.0, 5.0, 64, 4
    rng = np.random.default_rng(7)
    gen = t.Generator(device="cuda")
    gen.manual_seed(7)
    small = bench.batch_jobs(rng, nsmall)  # (N, S, -) per small job id
    gcfs = {S: O.config_gcf(S) for S in (8, 16, 32, 64)}
    d_gcf = {S: t.from_numpy(g).cuda() for S, g in gcfs.items()}
    pools = {jid: bench.gen_clustered_dev(t, 2_000_000, N, S, 800.0 * N / 8192, gen) + (N, S)
             for jid, (N, S, _) in enumerate(small)}
    big = bench.gen_clustered_dev(t, 10_000_000, 8192, 64, 800.0, gen)
    for k in range(nlarge):
        pools[1000 + k] = big + (8192, 64)
    t.cuda.synchronize()
    arr = np.cumsum(rng.exponential(1.0 / rate, int(rate * duration * 1.2) + 16))
    arr = arr[arr < duration]
    trace = []
    for when in arr:
        if rng.uniform() < 0.1:
            trace.append((float(when), 1000 + int(rng.integers(nlarge)), 10_000_000, 0))
        else:
            n = int(np.exp(rng.uniform(np.log(5e4), np.log(2e6))))
            trace.append((float(when), int(rng.integers(nsmall)), n,
                          int(rng.integers(0, 2_000_000 - n + 1))))
    # one more request on a job id of its own, in the middle of the trace
    solo = (duration / 2, 5000, 1_000_000, 123_456)
    pools[5000] = pools[0]
    trace.append(solo)
    trace.sort()
    print(f"[serve] {len(trace)} requests over {duration} s, "
          f"{sum(r[2] for r in trace)} visibilities", flush=True)

    def args(jid, n, off):
        uv, vals, N, S = pools[jid]
        return uv[off:off + n], vals[off:off + n], d_gcf[S], N, S

    import time
    with migrid.Server() as srv:
        tickets = []
        t0 = time.perf_counter()
        for when, jid, n, off in trace:
            now = time.perf_counter() - t0
            if when > now:
                time.sleep(when - now)
            tickets.append(srv.submit(jid, *args(jid, n, off)))
        for tk in tickets:
            srv.wait(tk)
        st = srv.stats()
        print(f"[serve] {st['completed']} completed in {st['batches']} batches, "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        assert st["completed"] == len(trace) and st["failed"] == 0
        grids = {jid: srv.read_grid(jid) for jid in {r[1] for r in trace}}
    # the large job ids: count x the oracle's grid of the 10M visibilities
    uv_b = big[0].cpu().numpy()
    vals_b = big[1].cpu().numpy()
    print(f"[serve] oracle (8192^2 / S=64 / 10M) on {threads()} threads", flush=True)
    ref_b, _ = O.grid_zz_omp(uv_b, vals_b, 8192, gcfs[64], 64, threads())
    for k in range(nlarge):
        cnt = sum(1 for r in trace if r[1] == 1000 + k)
        if cnt:
            e = rel_max_np(grids[1000 + k], cnt * ref_b)
            print(f"  job {1000 + k}: {cnt} requests, rel_max {e:.2e}", flush=True)
            assert e < 1e-5
    del ref_b, uv_b, vals_b
    # three small job ids of different supports: all their requests' visibilities
    load = {}
    for _, jid, n, _ in trace:
        if jid < nsmall:
            load[jid] = load.get(jid, 0) + n
    picked = []
    for S in (8, 32, 64):
        cands = sorted((v, jid) for jid, v in load.items() if small[jid][1] == S)
        if cands:
            picked.append(cands[0][1])
    for jid in picked:
        uv_p, vals_p, N, S = pools[jid]
        uv_p, vals_p = uv_p.cpu().numpy(), vals_p.cpu().numpy()
        sel = [(off, n) for _, j, n, off in trace if j == jid]
        uv_c = np.concatenate([uv_p[o:o + n] for o, n in sel])
        vals_c = np.concatenate([vals_p[o:o + n] for o, n in sel])
        ref, _ = O.grid_zz_omp(uv_c, vals_c, N, gcfs[S], S, threads())
        e = rel_max_np(grids[jid], ref)
        print(f"  job {jid} (N={N}, S={S}): {len(sel)} requests, rel_max {e:.2e}", flush=True)
        assert e < 1e-5
    # the solo request, alone on a fresh server and in the mix
    with migrid.Server() as srv:
        srv.wait(srv.submit(9, *args(5000, solo[2], solo[3])))
        alone = srv.read_grid(9)
    uv_s, vals_s, N, S = pools[5000]
    ref, _ = O.grid_zz_omp(uv_s[solo[3]:solo[3] + solo[2]].cpu().numpy(),
                           vals_s[solo[3]:solo[3] + solo[2]].cpu().numpy(), N, gcfs[S], S,
                           threads())
    e_mix, e_alone = rel_max_np(grids[5000], ref), rel_max_np(alone, ref)
    e_pair = O.rel_errors(grids[5000], alone)[0]
    print(f"  solo request: in the mix {e_mix:.2e}, alone {e_alone:.2e}, mix vs alone {e_pair:.2e}",
          flush=True)
    assert e_mix < 1e-5 and e_alone < 1e-5 and e_pair < 1e-5
