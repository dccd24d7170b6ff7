"""GPU tests of batched launches, hipGraph replay, the autotuner and the
tiled gridder's two coefficient paths (LDS-staged and direct from HBM).

FP64 results within 1e-12 relative max-norm of the CPU oracle; FP32 tiles
within 1e-5 of it.
"""
import numpy as np
import pytest

import migrid
import oracle as O

pytestmark = pytest.mark.gpu


def torch():
    import torch as t
    return t


def dev(a):
    return torch().from_numpy(np.ascontiguousarray(a)).cuda()


def cplx(g):
    g = g.cpu().numpy()
    return g[..., 0] + 1j * g[..., 1]


SHAPES = [(128, 8, 3000), (256, 16, 6000), (192, 32, 4000), (300, 64, 5000), (96, 10, 800)]


def make_jobs(dtype="f64", seed=0):
    t = torch()
    jobs, refs = [], []
    for i, (N, S, n) in enumerate(SHAPES):
        if i % 2:
            uv, vals = O.gen_clustered(n, N, S, mean_r=N / 8, seed=seed + i)
        else:
            uv, vals = O.gen_uniform(n, -S / 2, N + S / 2, seed=seed + i)
        gcf = O.config_gcf(S)
        refs.append(cplx_np(O.grid_zz(uv, vals, N, gcf, S)))
        jobs.append(dict(grid=t.empty((N, N, 2), dtype=t.float64 if dtype == "f64" else t.float32,
                                      device="cuda"),
                         uv=dev(uv), vals=dev(vals), gcf=dev(gcf), N=N, S=S))
    return jobs, refs


def cplx_np(p):
    return p[0] + 1j * p[1]


@pytest.mark.parametrize("arith,tol", [(migrid.ARITH_NATIVE, 1e-12), (migrid.ARITH_F32, 1e-5)])
def test_batch_matches_oracle(arith, tol):
    jobs, refs = make_jobs()
    ctx = migrid.Context()
    b = migrid.Batch(ctx, jobs, arith=arith)
    assert b.info()["njobs"] == len(jobs)
    b.run()
    torch().cuda.synchronize()
    for j, r in zip(jobs, refs):
        assert O.rel_errors(cplx(j["grid"]), r)[0] < tol, (j["N"], j["S"])


def test_graph_replays_match_eager_and_pick_up_new_data():
    t = torch()
    jobs, refs = make_jobs()
    ctx = migrid.Context()
    b = migrid.Batch(ctx, jobs, arith=migrid.ARITH_NATIVE)
    stream = t.cuda.Stream()
    other = make_jobs(seed=100)[0][0]
    with t.cuda.stream(stream):
        for rep in range(3):
            b.graph_launch(stream)
            stream.synchronize()
            for j, r in zip(jobs, refs):
                assert O.rel_errors(cplx(j["grid"]), r)[0] < 1e-12, (rep, j["N"])
            # eager work on the same context between replays
            ctx.grid(other["grid"], other["uv"], other["vals"], other["gcf"], other["N"], other["S"],
                     strategy="tiled", stream=stream)
            x = t.randn(1000, device="cuda").sum()
            stream.synchronize()
            assert t.isfinite(x)
        # new values in the same buffers: the replay grids them
        for j in jobs:
            j["vals"].mul_(-2.0)
        stream.synchronize()
        b.graph_launch(stream)
        stream.synchronize()
    for j, r in zip(jobs, refs):
        assert O.rel_errors(cplx(j["grid"]), -2.0 * r)[0] < 1e-12


def test_autotune_picks_a_listed_configuration_and_caches():
    jobs, refs = make_jobs()
    ctx = migrid.Context()
    best, ms = migrid.autotune(ctx, jobs, arith=migrid.ARITH_F32)
    assert best.tile in (16, 32, 64) and best.vis_per_wave in (64 * 64, 128 * 64, 256 * 64)
    assert best.wg_per_cu in (0, 1, 2) and best.stage_min in (0, 2, 16)  # LDS split
    assert ms > 0
    best2, ms2 = migrid.autotune(ctx, jobs, arith=migrid.ARITH_F32)
    assert (best2.tile, best2.vis_per_wave, best2.wg_per_cu, best2.stage_min, ms2) == \
        (best.tile, best.vis_per_wave, best.wg_per_cu, best.stage_min, ms)
    b = migrid.Batch(ctx, jobs, opts=best)
    b.run()
    torch().cuda.synchronize()
    for j, r in zip(jobs, refs):
        assert O.rel_errors(cplx(j["grid"]), r)[0] < 1e-5


@pytest.mark.parametrize("stage_min", [1, 1 << 30])
@pytest.mark.parametrize("arith,tol", [(migrid.ARITH_NATIVE, 1e-12), (migrid.ARITH_F32, 1e-5),
                                       (migrid.ARITH_F32_BF16GCF, 2e-2)])
def test_staged_and_direct_paths(stage_min, arith, tol):
    t = torch()
    N, S = 256, 64
    uv, vals = O.gen_clustered(20000, N, S, mean_r=30.0, seed=3)
    gcf = O.config_gcf(S)
    ref = cplx_np(O.grid_zz(uv, vals, N, gcf, S))
    grid = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    ctx = migrid.Context()
    for wg in (0, 1):
        # square tiles (tile option 32): the kernel with both coefficient
        # paths; the 128 x 64 tiles stage every sub-kernel
        ctx.grid(grid, dev(uv), dev(vals), dev(gcf), N, S, strategy="tiled", arith=arith,
                 stage_min=stage_min, wg_per_cu=wg, tile=32)
        t.cuda.synchronize()
        assert O.rel_errors(cplx(grid), ref)[0] < tol


@pytest.mark.parametrize("S", [8, 16, 32, 64])
def test_batch_of_one_support(S):
    """Jobs that share one support take the kernels specialised for it (tap
    kernel at S = 8 and 16, owner kernel at S = 32 and 64); grids of
    different sizes, a graph replay after a direct run."""
    t = torch()
    print(f"[batch one support] S={S}", flush=True)
    jobs, refs = [], []
    for i, (N, n) in enumerate([(96, 2000), (150, 4000), (64, 500)]):
        if i == 1:
            uv, vals = O.gen_clustered(n, N, S, mean_r=N / 8, seed=90 + i)
        else:
            uv, vals = O.gen_uniform(n, -S / 2, N + S / 2, seed=90 + i)
        gcf = O.config_gcf(S)
        refs.append(cplx_np(O.grid_zz(uv, vals, N, gcf, S)))
        jobs.append(dict(grid=t.empty((N, N, 2), dtype=t.float64, device="cuda"), uv=dev(uv),
                         vals=dev(vals), gcf=dev(gcf), N=N, S=S))
    ctx = migrid.Context()
    b = migrid.Batch(ctx, jobs)
    stream = t.cuda.Stream()
    for mode in ("run", "graph"):
        for j in jobs:
            j["grid"].fill_(5.0)
        t.cuda.synchronize()
        b.run(stream) if mode == "run" else b.graph_launch(stream)
        t.cuda.synchronize()
        for j, r in zip(jobs, refs):
            assert O.rel_errors(cplx(j["grid"]), r)[0] < 1e-12, (mode, j["N"])
    b.close()


@pytest.mark.parametrize("arith,tol", [(migrid.ARITH_NATIVE, 1e-12), (migrid.ARITH_F32, 1e-5)])
def test_batch_of_embedded_supports(arith, tol):
    """Owner jobs of supports 20, 24, 40, 48 and 64 in one launch: the first
    two run embedded in 32, the next two in 64, with one table shared by two
    jobs; graph replay into NaN-filled grids."""
    t = torch()
    print(f"[batch embedded supports] arith={arith}", flush=True)
    shapes = [(96, 20, 2000), (150, 24, 4000), (128, 40, 3000), (110, 48, 2500), (200, 64, 3000),
              (90, 24, 1500)]
    gcfs = {S: O.init_gcf(S) for S in (20, 24, 40, 48, 64)}
    dg = {S: dev(g) for S, g in gcfs.items()}
    jobs, refs = [], []
    for i, (N, S, n) in enumerate(shapes):
        uv, vals = O.gen_clustered(n, N, S, mean_r=N / 6, seed=70 + i)
        refs.append(cplx_np(O.grid_zz(uv, vals, N, gcfs[S], S)))
        jobs.append(dict(grid=t.empty((N, N, 2), dtype=t.float64, device="cuda"), uv=dev(uv),
                         vals=dev(vals), gcf=dg[S], N=N, S=S))
    ctx = migrid.Context()
    b = migrid.Batch(ctx, jobs, arith=arith)
    stream = t.cuda.Stream()
    for mode in ("run", "graph"):
        for j in jobs:
            j["grid"].fill_(float("nan"))
        t.cuda.synchronize()
        b.run(stream) if mode == "run" else b.graph_launch(stream)
        t.cuda.synchronize()
        for j, r in zip(jobs, refs):
            assert O.rel_errors(cplx(j["grid"]), r)[0] < tol, (mode, j["N"], j["S"])
    b.close()


@pytest.mark.parametrize("strategy", ["gather", "window"])
def test_batch_reproducible_mixed_supports(strategy):
    """GATHER / WINDOW over jobs of supports 24, 32, 48 and 64 in one batch:
    the support-32 jobs (24 embedded in 32) run in a second owner launch with
    a queue of their own, and split tiles of both launches share one ordered
    reduction.  FP64 within 1e-12 of the oracle, and bit for bit the same
    grids on a second run."""
    t = torch()
    print(f"[batch reproducible mixed] {strategy}", flush=True)
    shapes = [(160, 24, 6000), (200, 32, 8000), (140, 48, 5000), (220, 64, 9000)]
    jobs, refs = [], []
    for i, (N, S, n) in enumerate(shapes):
        uv, vals = O.gen_clustered(n, N, S, mean_r=N / 10, seed=120 + i)
        gcf = O.init_gcf(S)
        refs.append(cplx_np(O.grid_zz(uv, vals, N, gcf, S)))
        shape = (2, N, N) if strategy == "window" else (N, N, 2)
        jobs.append(dict(grid=t.empty(shape, dtype=t.float64, device="cuda"), uv=dev(uv),
                         vals=dev(vals), gcf=dev(gcf), N=N, S=S))
    ctx = migrid.Context()
    b = migrid.Batch(ctx, jobs, strategy=strategy, vis_per_wave=256)
    stream = t.cuda.Stream()
    runs = []
    for _ in range(2):
        for j in jobs:
            j["grid"].fill_(float("nan"))
        t.cuda.synchronize()
        b.run(stream)
        t.cuda.synchronize()
        runs.append([j["grid"].cpu().numpy().copy() for j in jobs])
    for j, r, g0, g1 in zip(jobs, refs, runs[0], runs[1]):
        np.testing.assert_array_equal(g0, g1)
        got = g0[0] + 1j * g0[1] if strategy == "window" else g0[..., 0] + 1j * g0[..., 1]
        assert O.rel_errors(got, r)[0] < 1e-12, (strategy, j["S"])
    b.close()


@pytest.mark.parametrize("strategy", ["tiled", "gather"])
def test_batch_coverage_mixed_supports(strategy):
    """Coverage counts (constant taps, so no support embedding) for supports
    24, 32 and 64 in one batch: the support-32 jobs run in a second owner
    launch with their own queue, the others in the general instantiation;
    counts bit-exact against the oracle, GATHER's split tiles summed through
    one ordered reduction across both launches."""
    t = torch()
    print(f"[batch coverage mixed] {strategy}", flush=True)
    shapes = [(120, 24, 3000), (150, 32, 4000), (180, 64, 4000), (100, 32, 2000)]
    jobs, refs = [], []
    for i, (N, S, n) in enumerate(shapes):
        uv, vals = O.gen_clustered(n, N, S, mean_r=N / 8, seed=140 + i)
        gcf = O.config_gcf(S)
        refs.append(O.grid_zz(uv, vals, N, gcf, S, debug1=True))
        jobs.append(dict(grid=t.empty((N, N, 2), dtype=t.float64, device="cuda"), uv=dev(uv),
                         vals=dev(vals), gcf=dev(gcf), N=N, S=S))
    ctx = migrid.Context()
    b = migrid.Batch(ctx, jobs, strategy=strategy, debug_coverage=True, vis_per_wave=128,
                     layout=migrid.LAYOUT_INTERLEAVED)
    for j in jobs:
        j["grid"].fill_(float("nan"))
    t.cuda.synchronize()
    b.run()
    t.cuda.synchronize()
    for j, r in zip(jobs, refs):
        g = j["grid"].cpu().numpy()
        np.testing.assert_array_equal(g[..., 0], r[0])
        np.testing.assert_array_equal(g[..., 1], r[1])
    b.close()
