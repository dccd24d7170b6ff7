"""GPU tests of the grid <-> image transform (migrid_ctx_fft) against the
oracle's numpy restatement (itself pinned by the DFT definition): double
within 1e-11 and float within 2e-5 of the largest magnitude; the round trip
and Parseval's identity at 8192^2; the gridding -> image chain."""
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


def c2(a):
    """complex (N, N) -> float (N, N, 2)"""
    return np.stack([a.real, a.imag], -1)


def cplx(a):
    a = a.cpu().numpy().astype(np.float64) if hasattr(a, "cpu") else a
    return a[..., 0] + 1j * a[..., 1]


def close(got, want, tol):
    return np.abs(got - want).max() <= tol * np.abs(want).max()


@pytest.mark.parametrize("N", [64, 100, 256])
@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_fft_matches_oracle(N, dt):
    t = torch()
    rng = np.random.default_rng(N)
    g = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    taper = 0.5 + rng.random(N)
    tdt, tol = (t.float64, 1e-11) if dt == "f64" else (t.float32, 2e-5)
    ctx = migrid.Context()
    x = dev(c2(g)).to(tdt)
    d_taper = dev(taper)
    x0 = x.clone()
    got = cplx(ctx.fft(x))
    assert close(got, O.grid_to_image(g), tol)
    assert t.equal(x, x0)  # an out-of-place transform leaves its input alone
    got = cplx(ctx.fft(x, scale=0.25, taper=d_taper))
    assert close(got, O.grid_to_image(g, 0.25, taper), tol)
    got = cplx(ctx.fft(x, direction=migrid.FFT_TO_GRID, scale=1.5, taper=d_taper))
    assert close(got, O.image_to_grid(g, 1.5, taper), tol)
    # planar input
    xp = dev(np.stack([g.real, g.imag])).to(tdt)
    got = cplx(ctx.fft(xp, planar=True))
    assert close(got, O.grid_to_image(g), tol)
    # in place (interleaved)
    y = x.clone()
    r = ctx.fft(y, out=y)
    assert r.data_ptr() == y.data_ptr()
    assert close(cplx(y), O.grid_to_image(g), tol)


def test_fft_rejects_bad_arguments():
    t = torch()
    ctx = migrid.Context()
    with pytest.raises(migrid.MigridError):
        ctx.fft(t.zeros((7, 7, 2), dtype=t.float64, device="cuda"))  # odd size
    xp = t.zeros((2, 8, 8), dtype=t.float64, device="cuda")
    with pytest.raises(migrid.MigridError):
        ctx.fft(xp, out=xp, planar=True)  # planar in place


def test_fft_round_trip_and_parseval_at_8192():
    t = torch()
    N = 8192
    print("[8192^2] fft round trip", flush=True)
    ctx = migrid.Context()
    g = t.randn((N, N, 2), dtype=t.float64, device="cuda",
                generator=t.Generator(device="cuda").manual_seed(11))
    img = ctx.fft(g)
    e_grid = float((g * g).sum())
    e_img = float((img * img).sum())
    assert abs(e_img - N * N * e_grid) <= 1e-10 * N * N * e_grid  # Parseval, unnormalised
    back = ctx.fft(img, direction=migrid.FFT_TO_GRID, scale=1.0 / (N * N))
    err = float((back - g).abs().max())
    assert err <= 1e-12 * float(g.abs().max())
    # float
    g32 = g.float()
    back32 = ctx.fft(ctx.fft(g32), direction=migrid.FFT_TO_GRID, scale=1.0 / (N * N))
    assert float((back32 - g32).abs().max()) <= 1e-5 * float(g32.abs().max())


def test_gridding_then_image_matches_oracle():
    """The chain a serving loop runs: grid visibilities, then the image."""
    N, S, n = 256, 16, 20000
    uv, vals = O.gen_clustered(n, N, S, mean_r=30.0, seed=5)
    gcf = O.config_gcf(S)
    want_grid = O.grid_zz(uv, vals, N, gcf, S)
    want = O.grid_to_image(want_grid[0] + 1j * want_grid[1])
    t = torch()
    ctx = migrid.Context()
    grid = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    ctx.grid(grid, dev(uv), dev(vals), dev(gcf), N, S, strategy="tiled")
    got = cplx(ctx.fft(grid))
    assert close(got, want, 1e-11)
