"""Seeded random sweep of shapes against the oracle.

Grid edges from 3 to 700 (odd sizes, sizes below the support), even supports
2..64, visibility counts from 0 up, windows crossing every edge or a dense
cluster that piles thousands of visibilities into one tile and sub-kernel;
each case through every gridding strategy (FP64, 1e-12 relative max-norm),
FP32 tiles (1e-5) and BF16-GCF tiles (1e-5 of the oracle on the bf16-rounded
table), accumulation into a filled grid, the planar layout, a tile edge drawn
per case; and the degridder on the same shapes (1e-12 / 1e-5).  A reference
that is all zeros (no visibility, or every tap off the grid) must be matched
exactly.
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


def make_cases(n=24, seed=20261203):
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(n):
        N = int(rng.choice([3, 5, 16, 17, 31, 64, 100, 129, 257, 700, int(rng.integers(6, 400))]))
        S = int(2 * rng.integers(1, 33))
        npts = int(rng.choice([0, 1, 2, 17, 1000, 5000]))
        dense = bool(rng.random() < 0.3) and npts > 0
        tile = int(rng.choice([16, 32, 64]))
        cases.append((i, N, S, npts, dense, tile))
    return cases


CASES = make_cases()


def inputs(i, N, S, npts, dense):
    rng = np.random.default_rng(1000 + i)
    if dense:
        c = rng.uniform(0, N, size=2)
        uv = c + rng.normal(scale=2.0, size=(npts, 2))
    else:
        uv = rng.uniform(-S, N + S, size=(npts, 2))
    uv = uv.astype(np.float32).astype(np.float64)  # float-representable, as the reference's floorf assumes
    vals = rng.normal(size=(npts, 2))
    return uv, vals


def bf16_round(a):
    f = np.ascontiguousarray(a, np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def cplx_grid(t, planar):
    g = t.cpu().numpy().astype(np.float64)
    return g[0] + 1j * g[1] if planar else g[..., 0] + 1j * g[..., 1]


def check(got, want, tol, what):
    m = np.abs(want).max() if want.size else 0.0
    if m == 0.0:
        assert np.abs(got).max(initial=0.0) == 0.0, what
        return
    err = np.abs(got - want).max() / m
    assert err <= tol, f"{what}: rel max {err:.3e} > {tol}"


@pytest.mark.parametrize("case", CASES, ids=lambda c: "i%d_N%d_S%d_n%d%s_T%d" % (
    c[0], c[1], c[2], c[3], "_dense" if c[4] else "", c[5]))
def test_gridding_sweep(case):
    i, N, S, npts, dense, tile = case
    t = torch()
    print(f"[sweep] case {case}", flush=True)
    uv, vals = inputs(i, N, S, npts, dense)
    gcf = O.config_gcf(S)
    ref = O.grid_zz(uv, vals, N, gcf, S)
    want = ref[0] + 1j * ref[1]
    ctx = migrid.Context()
    d_uv, d_vals, d_gcf = dev(uv), dev(vals), dev(gcf)
    for strategy in ("scatter", "tiled", "gather", "window"):
        if strategy != "scatter" and S > 64:
            continue
        planar = strategy == "window"
        grid = t.empty((2, N, N) if planar else (N, N, 2), dtype=t.float64, device="cuda")
        kw = dict(strategy=strategy)
        if strategy != "scatter":
            kw["tile"] = tile
        ctx.grid(grid, d_uv, d_vals, d_gcf, N, S, **kw)
        check(cplx_grid(grid, planar), want, 1e-12, f"{strategy} fp64")
    # FP32 tiles, planar, accumulating into a filled grid
    rng = np.random.default_rng(7 + i)
    base = rng.normal(size=(2, N, N))
    grid = dev(base)
    ctx.grid(grid, d_uv, d_vals, d_gcf, N, S, strategy="tiled", tile=tile, arith=migrid.ARITH_F32,
             layout=migrid.LAYOUT_PLANAR, accumulate=True)
    got = cplx_grid(grid, True) - (base[0] + 1j * base[1])
    # the base values are exact in the grid; the FP32 error scales with the gridded part
    m = max(np.abs(want).max(initial=0.0), 1e-300)
    assert np.abs(got - want).max(initial=0.0) <= 1e-5 * m + 1e-12 * np.abs(base).max(), "fp32 accumulate"
    # BF16 GCF tiles against the oracle on the bf16-rounded table
    gb = bf16_round(gcf)
    refb = O.grid_zz(uv, vals, N, gb, S)
    grid = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    ctx.grid(grid, d_uv, d_vals, d_gcf, N, S, strategy="tiled", tile=tile,
             arith=migrid.ARITH_F32_BF16GCF)
    check(cplx_grid(grid, False), refb[0] + 1j * refb[1], 1e-5, "bf16 gcf")


@pytest.mark.parametrize("case", CASES, ids=lambda c: "i%d_N%d_S%d_n%d%s" % (
    c[0], c[1], c[2], c[3], "_dense" if c[4] else ""))
def test_degridding_sweep(case):
    i, N, S, npts, dense, tile = case
    print(f"[sweep degrid] case {case}", flush=True)
    uv, _ = inputs(i, N, S, npts, dense)
    gcf = O.config_gcf(S)
    rng = np.random.default_rng(50 + i)
    g = rng.normal(size=(2, N, N))
    want = O.degrid_zz(uv, g, N, gcf, S)
    want = want[:, 0] + 1j * want[:, 1]
    ctx = migrid.Context()
    grid = dev(np.moveaxis(g, 0, -1))
    for arith, tol in ((migrid.ARITH_NATIVE, 1e-12), (migrid.ARITH_F32, 1e-5)):
        out = ctx.degrid(grid, dev(uv), dev(gcf), N, S, arith=arith).cpu().numpy()
        check(out[:, 0] + 1j * out[:, 1], want, tol, f"degrid arith {arith}")
