"""GPU tests of the input side: (u, v, w) -> pixels (bit-exact against the
CPU restatement) and gridding with a GCF table per visibility (against the
oracle's gridCPU summed over the tables: FP64 within 1e-12, FP32 within
1e-5)."""
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


def baselines(nrow, seed, scale):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, (nrow, 3))


def test_uvw_to_pixels_matches_the_restatement_bitwise():
    t = torch()
    uvw = baselines(5000, 1, 800.0)
    f = np.linspace(1.0e8, 1.5e8, 7)
    want_uv, want_g = O.uvw_to_pixels(uvw, f, 2048, 2e-5, nplanes=8, w_max=150.0)
    uv, g = migrid.uvw_to_pixels(dev(uvw), dev(f), 2048, 2e-5, nplanes=8, w_max=150.0)
    t.cuda.synchronize()
    np.testing.assert_array_equal(uv.cpu().numpy(), want_uv)
    np.testing.assert_array_equal(g.cpu().numpy(), want_g)
    uv2, g2 = migrid.uvw_to_pixels(dev(uvw), dev(f), 2048, 2e-5)
    assert g2 is None
    np.testing.assert_array_equal(uv2.cpu().numpy(), want_uv)


def gcf_set(S, ng):
    # w-plane tables: the Gaussian x chirp with a chirp rate per plane
    return np.stack([O.config_gcf(S, beta=0.02 + 0.03 * g) for g in range(ng)])


def oracle_indexed(uv, vals, gidx, gcfs, N, S):
    out = np.zeros((2, N, N))
    for g in range(gcfs.shape[0]):
        m = gidx == g
        if m.any():
            out += O.grid_zz(uv[m], vals[m], N, gcfs[g], S)
    return out[0] + 1j * out[1]


@pytest.mark.parametrize("strategy,arith,tol", [("tiled", migrid.ARITH_NATIVE, 1e-12),
                                                ("tiled", migrid.ARITH_F32, 1e-5),
                                                ("gather", migrid.ARITH_NATIVE, 1e-12)])
def test_indexed_gridding_matches_oracle(strategy, arith, tol):
    t = torch()
    N, S, ng = 512, 16, 5
    uvw = baselines(3000, 2, 300.0)
    f = np.array([1.0e8, 1.2e8, 1.4e8])
    uv, g = migrid.uvw_to_pixels(dev(uvw), dev(f), N, 4e-5, nplanes=ng, w_max=120.0)
    t.cuda.synchronize()
    uv_h, g_h = uv.cpu().numpy(), g.cpu().numpy()
    keep = ((np.floor(uv_h) - S // 2 + 1 >= 0) & (np.floor(uv_h) + S // 2 <= N - 1)).all(1)
    uv_h, g_h = uv_h[keep], g_h[keep]
    assert len(np.unique(g_h)) > 2
    rng = np.random.default_rng(3)
    vals = rng.normal(size=(uv_h.shape[0], 2))
    gcfs = gcf_set(S, ng)
    want = oracle_indexed(uv_h, vals, g_h, gcfs, N, S)
    grid = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    ctx = migrid.Context()
    ctx.grid_indexed(grid, dev(uv_h), dev(vals), dev(gcfs), dev(g_h), N, S, strategy=strategy,
                     arith=arith)
    t.cuda.synchronize()
    got = grid.cpu().numpy()
    assert O.rel_errors(got[..., 0] + 1j * got[..., 1], want)[0] < tol


def test_indexed_gridding_edge_cases():
    t = torch()
    N, S = 128, 8
    uv, vals = O.gen_uniform(2000, -S, N + S, seed=9)
    gcfs = gcf_set(S, 3)
    ctx = migrid.Context()
    grid = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    # one table for everyone = plain tiled gridding
    ctx.grid_indexed(grid, dev(uv), dev(vals), dev(gcfs[1]), dev(np.zeros(2000, np.int32)), N, S)
    t.cuda.synchronize()
    want = O.grid_zz(uv, vals, N, gcfs[1], S)
    got = grid.cpu().numpy()
    assert O.rel_errors(got[..., 0] + 1j * got[..., 1], want[0] + 1j * want[1])[0] < 1e-12
    # accumulate onto what is there
    gi = (np.arange(2000) % 3).astype(np.int32)
    ctx.grid_indexed(grid, dev(uv), dev(vals), dev(gcfs), dev(gi), N, S, accumulate=True)
    t.cuda.synchronize()
    want2 = want[0] + 1j * want[1] + oracle_indexed(uv, vals, gi, gcfs, N, S)
    got = grid.cpu().numpy()
    assert O.rel_errors(got[..., 0] + 1j * got[..., 1], want2)[0] < 1e-12
    # an index out of range is refused, naming the count
    bad = gi.copy()
    bad[7] = 3
    with pytest.raises(migrid.MigridError) as e:
        ctx.grid_indexed(grid, dev(uv), dev(vals), dev(gcfs), dev(bad), N, S)
    assert e.value.code == -1 and "1 visibilities" in str(e.value)
