"""Parity at the benchmark's full sizes.

2048^2 grid, S = 32, 1M uniform visibilities (the reference's CPU-runnable
case): every strategy against the oracle's gridCPU (row-band OpenMP form,
bit-identical to the serial loop), FP64 within 1e-12, FP32 tiles within 1e-5.

8192^2 grid, S = 64, 10M radially clustered visibilities: too big for the
oracle in a test, so size-independent properties: every FP64 strategy agrees
with the plain FP64 scatter within 1e-12 and FP32 tiles within 1e-5; the
coverage map (DEBUG1: +1 per tap) is bit-identical between the scatter and
the tiled kernel and sums to exactly npts * S^2; gridding is linear
(values x 2 -> grid x 2, exact up to summation order).
"""
import os

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


def rel_max(a, b):
    t = torch()
    d = t.linalg.vector_norm((a - b).reshape(-1, 2) if a.shape[-1] == 2 else (a - b).reshape(2, -1).T, dim=-1)
    m = t.linalg.vector_norm(b.reshape(-1, 2) if b.shape[-1] == 2 else b.reshape(2, -1).T, dim=-1)
    return float(d.max() / m.max())


def threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def test_2048_s32_1m_matches_gridcpu():
    t = torch()
    N, S, n = 2048, 32, 1_000_000
    print(f"[2048^2 S=32 1M] oracle with {threads()} threads", flush=True)
    uv, vals = O.gen_uniform(n, 64, 1984, seed=1541617)
    gcf = O.config_gcf(S)
    ref, _ = O.grid_zz_omp(uv, vals, N, gcf, S, threads())
    ref_t = t.from_numpy(np.ascontiguousarray(np.moveaxis(ref, 0, -1))).cuda()
    d_uv, d_vals, d_gcf = dev(uv), dev(vals), dev(gcf)
    grid = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    ctx = migrid.Context()
    for strategy, kw, tol in (("scatter", {}, 1e-12), ("tiled", {}, 1e-12), ("gather", {}, 1e-12),
                              ("tiled", dict(arith=migrid.ARITH_F32), 1e-5),
                              ("tiled", dict(arith=migrid.ARITH_F32_BF16GCF), 2e-2)):
        ctx.grid(grid, d_uv, d_vals, d_gcf, N, S, strategy=strategy, layout=migrid.LAYOUT_INTERLEAVED,
                 **kw)
        t.cuda.synchronize()
        e = rel_max(grid, ref_t)
        print(f"  {strategy} {kw}: rel_max {e:.3e}", flush=True)
        assert e < tol, (strategy, kw, e)


@pytest.fixture(scope="module")
def big():
    t = torch()
    N, S, n = 8192, 64, 10_000_000
    uv, vals = O.gen_clustered(n, N, S, mean_r=800.0, seed=1541618)
    d = dict(N=N, S=S, n=n, uv=dev(uv), vals=dev(vals), gcf=dev(O.config_gcf(S)))
    del uv, vals
    ctx = migrid.Context()
    ref = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    ctx.grid(ref, d["uv"], d["vals"], d["gcf"], N, S, strategy="scatter")
    t.cuda.synchronize()
    d.update(ctx=ctx, ref=ref)
    yield d
    del d
    t.cuda.empty_cache()


@pytest.mark.parametrize("strategy,kw,tol", [
    ("tiled", {}, 1e-12), ("gather", {}, 1e-12),
    ("window", dict(layout=migrid.LAYOUT_INTERLEAVED), 1e-12),
    ("tiled", dict(arith=migrid.ARITH_F32), 1e-5),
])
def test_8192_s64_10m_strategies_agree(big, strategy, kw, tol):
    t = torch()
    N, S = big["N"], big["S"]
    print(f"[8192^2 S=64 10M] {strategy} {kw}", flush=True)
    g = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    big["ctx"].grid(g, big["uv"], big["vals"], big["gcf"], N, S, strategy=strategy, **kw)
    t.cuda.synchronize()
    e = rel_max(g, big["ref"])
    print(f"  rel_max vs scatter {e:.3e}", flush=True)
    assert e < tol


def test_8192_s64_10m_coverage_bit_exact_and_linear(big):
    t = torch()
    N, S, n = big["N"], big["S"], big["n"]
    ctx = big["ctx"]
    print("[8192^2 S=64 10M] coverage + linearity", flush=True)
    a = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    b = t.empty_like(a)
    ctx.grid(a, big["uv"], big["vals"], big["gcf"], N, S, strategy="scatter", debug_coverage=True)
    ctx.grid(b, big["uv"], big["vals"], big["gcf"], N, S, strategy="tiled", debug_coverage=True)
    t.cuda.synchronize()
    assert t.equal(a, b)
    assert float(a[..., 0].sum()) == float(n) * S * S
    del a
    v2 = big["vals"] * 2.0
    ctx.grid(b, big["uv"], v2, big["gcf"], N, S, strategy="tiled")
    t.cuda.synchronize()
    assert rel_max(b, 2.0 * big["ref"]) < 1e-12
