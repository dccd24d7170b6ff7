"""GPU tests of degridding (the adjoint of gridding) against the CPU
restatement oracle_degrid_zz (itself pinned by the adjoint identity):
double products within 1e-12 relative max-norm, FP32 products within 1e-5;
and the adjoint identity <grid(v), g> = <v, degrid(g)> between the GPU
gridder and the GPU degridder at full size (8192^2, S = 64)."""
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


def cplx(a):
    return a[..., 0] + 1j * a[..., 1]


@pytest.mark.parametrize("S,N,npts", [(8, 64, 3000), (16, 100, 4000), (32, 200, 5000),
                                      (64, 300, 3000), (6, 70, 1000), (64, 96, 500)])
def test_degrid_matches_oracle(S, N, npts):
    rng = np.random.default_rng(S * 7 + N)
    uv, _ = O.gen_uniform(npts, -S, N + S, seed=S + N)   # windows cross every edge
    gcf = O.config_gcf(S)
    g = rng.normal(size=(2, N, N))
    want = cplx(O.degrid_zz(uv, g, N, gcf, S))
    ctx = migrid.Context()
    grid = dev(np.moveaxis(g, 0, -1))
    got = cplx(ctx.degrid(grid, dev(uv), dev(gcf), N, S).cpu().numpy())
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    got32 = cplx(ctx.degrid(grid, dev(uv), dev(gcf), N, S, arith=migrid.ARITH_F32).cpu().numpy())
    assert np.abs(got32 - want).max() <= 1e-5 * np.abs(want).max()
    # planar grid layout gives the same values
    gp = dev(g)
    got_p = cplx(ctx.degrid(gp, dev(uv), dev(gcf), N, S, layout=migrid.LAYOUT_PLANAR).cpu().numpy())
    assert np.abs(got_p - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("tile", [16, 32, 64])
@pytest.mark.parametrize("vpw", [1, 2, 4])
def test_degrid_tiles_and_groups(tile, vpw):
    """Every tile edge and visibilities-per-pass choice; a dense centre puts
    thousands of visibilities in one tile (split into several work items) and
    long runs of one sub-kernel (groups sharing the GCF rows)."""
    N, S = 160, 32
    rng = np.random.default_rng(tile + vpw)
    uv, _ = O.gen_uniform(6000, -S, N + S, seed=tile * 10 + vpw)
    uv[:4000] = N / 2 + rng.normal(scale=6.0, size=(4000, 2))
    gcf = O.config_gcf(S)
    g = rng.normal(size=(2, N, N))
    want = cplx(O.degrid_zz(uv, g, N, gcf, S))
    ctx = migrid.Context()
    grid = dev(np.moveaxis(g, 0, -1))
    for arith, tol in ((migrid.ARITH_NATIVE, 1e-12), (migrid.ARITH_F32, 1e-5)):
        got = cplx(ctx.degrid(grid, dev(uv), dev(gcf), N, S, arith=arith, tile=tile,
                              vis_per_wave=vpw).cpu().numpy())
        assert np.abs(got - want).max() <= tol * np.abs(want).max()


def test_degrid_float_types_and_empty():
    rng = np.random.default_rng(5)
    N, S = 128, 16
    uv, _ = O.gen_uniform(2000, 0, N, seed=5)
    gcf = O.config_gcf(S)
    g = rng.normal(size=(2, N, N))
    want = cplx(O.degrid_zz(uv.astype(np.float32).astype(np.float64), g, N,
                            gcf.astype(np.float32).astype(np.float64), S))
    ctx = migrid.Context()
    t = torch()
    for grid_dt in (t.float64, t.float32):
        grid = dev(np.moveaxis(g, 0, -1)).to(grid_dt)
        got = cplx(ctx.degrid(grid, dev(uv.astype(np.float32)), dev(gcf.astype(np.float32)), N, S)
                   .cpu().numpy().astype(np.float64))
        assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()
    out = ctx.degrid(dev(np.moveaxis(g, 0, -1)), dev(np.zeros((0, 2))), dev(gcf), N, S)
    assert out.shape == (0, 2)


def test_adjoint_identity_at_full_size():
    t = torch()
    N, S, n = 8192, 64, 2_000_000
    print("[8192^2 S=64 2M] adjoint identity", flush=True)
    uv, v = O.gen_clustered(n, N, S, mean_r=800.0, seed=77)
    gcf = O.config_gcf(S)
    ctx = migrid.Context()
    d_uv, d_v, d_gcf = dev(uv), dev(v), dev(gcf)
    Av = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    ctx.grid(Av, d_uv, d_v, d_gcf, N, S, strategy="tiled")
    g = t.randn((N, N, 2), dtype=t.float64, device="cuda",
                generator=t.Generator(device="cuda").manual_seed(3))
    Ag = ctx.degrid(g, d_uv, d_gcf, N, S)
    t.cuda.synchronize()
    lhs = t.view_as_complex(Av).conj().mul(t.view_as_complex(g)).sum()
    rhs = t.view_as_complex(d_v.contiguous()).conj().mul(t.view_as_complex(Ag)).sum()
    assert abs(complex(lhs) - complex(rhs)) <= 1e-11 * abs(complex(lhs))
