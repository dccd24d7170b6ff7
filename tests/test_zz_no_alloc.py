"""migrid_grid_noalloc: the allocation-free entry point, against
migrid_ctx_grid on the same inputs.

Two calls in a row share the library's static scratch array, the second on
a smaller job than the first, for every strategy: FP64 within 1e-12
relative max-norm of the context path (split tiles flush with atomics, so
summation order may differ), gather and window bit for bit (their split
tiles are summed in a fixed order), FP32 tiles within 1e-5.  Outputs are
filled with NaN before each call.  A job beyond the array's size is refused
before any launch, leaving the output untouched.
"""
import numpy as np
import pytest

import migrid
import oracle as O

pytestmark = pytest.mark.gpu

NAN = float("nan")


def torch():
    import torch as t
    return t


def dev(a):
    return torch().from_numpy(np.ascontiguousarray(a)).cuda()


def rel_max(a, b):
    t = torch()
    a = a.double().reshape(-1, 2) if a.shape[-1] == 2 else a.double().reshape(2, -1).T
    b = b.double().reshape(-1, 2) if b.shape[-1] == 2 else b.double().reshape(2, -1).T
    return float(t.linalg.vector_norm(a - b, dim=-1).max() / t.linalg.vector_norm(b, dim=-1).max())


CASES = [  # strategy, options, tolerance (None: bit for bit)
    ("tiled", {}, 1e-12),
    ("tiled", dict(arith=migrid.ARITH_F32), 1e-5),
    ("gather", {}, None),
    ("window", {}, None),
    ("scatter", {}, 1e-12),
    ("tiled", dict(gcf_mode=migrid.GCF_COMPAT), 1e-12),
]

# (N, S, visibilities, mean radius): the second job is smaller than the first
JOBS = [(700, 32, 300_000, 60.0), (256, 16, 20_000, 30.0)]


@pytest.mark.parametrize("strategy,kw,tol", CASES)
def test_noalloc_matches_context_path_twice(strategy, kw, tol):
    t = torch()
    ctx = migrid.Context()
    for N, S, n, r in JOBS:
        print(f"[noalloc] {strategy} {kw} N={N} S={S} n={n}", flush=True)
        uv, vals = O.gen_clustered(n, N, S, mean_r=r, seed=N + S)
        d_uv, d_vals, d_gcf = dev(uv), dev(vals), dev(O.config_gcf(S))
        planar = strategy == "window"
        shape = (2, N, N) if planar else (N, N, 2)
        want = t.full(shape, NAN, dtype=t.float64, device="cuda")
        got = t.full(shape, NAN, dtype=t.float64, device="cuda")
        ctx.grid(want, d_uv, d_vals, d_gcf, N, S, strategy=strategy, **kw)
        migrid.grid_noalloc(got, d_uv, d_vals, d_gcf, N, S, strategy=strategy, **kw)
        t.cuda.synchronize()
        assert not t.isnan(got).any()
        if tol is None:
            assert t.equal(got, want)
        else:
            assert rel_max(got, want) < tol
        if strategy != "scatter" and not kw:
            ref = O.grid_zz(uv, vals, N, O.config_gcf(S), S)
            ref_t = t.from_numpy(ref if planar else np.ascontiguousarray(np.moveaxis(ref, 0, -1))).cuda()
            assert rel_max(got, ref_t) < 1e-12


def test_noalloc_accumulates_and_refuses_oversized_jobs():
    t = torch()
    N, S = 300, 32
    uv, vals = O.gen_clustered(50_000, N, S, mean_r=40.0, seed=5)
    d_uv, d_vals, d_gcf = dev(uv), dev(vals), dev(O.config_gcf(S))
    g = t.zeros((N, N, 2), dtype=t.float64, device="cuda")
    migrid.grid_noalloc(g, d_uv, d_vals, d_gcf, N, S, strategy="tiled")
    migrid.grid_noalloc(g, d_uv, d_vals, d_gcf, N, S, strategy="tiled", accumulate=True)
    t.cuda.synchronize()
    ref = O.grid_zz(uv, vals, N, O.config_gcf(S), S)
    ref_t = t.from_numpy(np.ascontiguousarray(np.moveaxis(2.0 * ref, 0, -1))).cuda()
    assert rel_max(g, ref_t) < 1e-12
    # 10M visibilities at 8192^2 / S = 64 need far more than the static array
    big_uv = t.full((10_000_000, 2), 4096.0, dtype=t.float64, device="cuda")
    big_g = t.full((8192, 8192, 2), NAN, dtype=t.float64, device="cuda")
    with pytest.raises(migrid.MigridError) as e:
        migrid.grid_noalloc(big_g, big_uv, big_uv, dev(O.config_gcf(64)), 8192, 64,
                            strategy="tiled")
    assert e.value.code == -4  # MIGRID_ENOTSUP
    assert "scratch" in str(e.value)
    t.cuda.synchronize()
    assert t.isnan(big_g).all()
