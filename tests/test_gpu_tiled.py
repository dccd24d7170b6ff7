"""GPU parity of the tiled owner gridder (TILED, GATHER, WINDOW strategies).

Tolerances: FP64 (native) within 1e-12 relative max-norm of the oracle;
coverage counts bit-exact; FP32 tiles within 1e-5 of the FP64 oracle; BF16
GCF tiles within 1e-5 of the oracle run on the bf16-rounded table (the error
against the exact table, about 1e-3, is reported by the benchmark, not
gated); float pairs within float rounding.
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


def run(uv, vals, gcf, N, S, out_dtype="f64", planar=False, grid_lo=False, **kw):
    t = torch()
    dt = t.float64 if out_dtype == "f64" else t.float32
    grid = t.empty((2, N, N) if planar else (N, N, 2), dtype=dt, device="cuda")
    lo = t.zeros_like(grid) if grid_lo else None
    ctx = migrid.Context()
    ctx.grid(grid, dev(uv), dev(vals), dev(gcf) if gcf is not None else None, N, S,
             grid_lo=lo, **kw)
    t.cuda.synchronize()
    g = grid.cpu().numpy().astype(np.float64)
    if lo is not None:
        g = g + lo.cpu().numpy().astype(np.float64)
    if planar:
        return g[0] + 1j * g[1]
    return g[..., 0] + 1j * g[..., 1]


def cplx(p):
    return p[0] + 1j * p[1]


STRATS = ["tiled", "gather", "window"]


@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("S,N,npts", [(8, 100, 4000), (16, 128, 4000), (32, 200, 6000), (64, 300, 5000),
                                      (6, 70, 1000), (10, 90, 1000), (48, 160, 2000)])
def test_native_f64_matches_gridcpu(strategy, S, N, npts):
    uv, vals = O.gen_uniform(npts, -S / 2, N + S / 2, seed=S * 7 + N)
    gcf = O.init_gcf(S)
    ref = O.grid_zz(uv, vals, N, gcf, S)
    planar = strategy == "window"
    g = run(uv, vals, gcf, N, S, planar=planar, strategy=strategy)
    assert O.rel_errors(g, cplx(ref))[0] < 1e-12


@pytest.mark.parametrize("strategy", STRATS)
def test_coverage_bit_exact(strategy):
    N, S = 150, 32
    uv, vals = O.gen_uniform(3000, -20, N + 20, seed=31)
    ref = O.grid_zz(uv, vals, N, O.config_gcf(S), S, debug1=True)
    g = run(uv, vals, O.config_gcf(S), N, S, planar=strategy == "window", strategy=strategy,
            debug_coverage=True)
    np.testing.assert_array_equal(g.real, ref[0])
    np.testing.assert_array_equal(g.imag, ref[1])


def test_split_tiles_use_atomic_flush():
    # a dense cluster and a tiny item target: many items per tile
    N, S = 256, 64
    uv, vals = O.gen_clustered(40000, N, S, mean_r=12.0, seed=5)
    gcf = O.config_gcf(S)
    ref = cplx(O.grid_zz(uv, vals, N, gcf, S))
    for vpw in (16, 300, 0):
        g = run(uv, vals, gcf, N, S, strategy="tiled", vis_per_wave=vpw)
        assert O.rel_errors(g, ref)[0] < 1e-12, vpw
    g32 = run(uv, vals, gcf, N, S, strategy="tiled", tile=32)
    assert O.rel_errors(g32, ref)[0] < 1e-12


@pytest.mark.parametrize("pair,tol", [("zc", 2e-6), ("cc", 2e-5)])
@pytest.mark.parametrize("strategy", STRATS)
def test_float_pairs(strategy, pair, tol):
    N, S = 128, 16
    uv, vals = O.gen_uniform(4000, 0, N, seed=9)
    gcf = O.config_gcf(S)
    ref = O.grid_zc(uv, vals, N, gcf, S) if pair == "zc" else O.grid_cc(uv, vals, N, gcf, S)
    g = run(uv.astype(np.float32), vals.astype(np.float32), gcf.astype(np.float32), N, S,
            out_dtype="f64" if pair == "zc" else "f32", planar=strategy == "window", strategy=strategy)
    assert O.rel_errors(g, cplx(ref))[0] < tol


@pytest.mark.parametrize("strategy,fixed", [("tiled", False), ("gather", True), ("window", True)])
def test_compute_gcf_compat(strategy, fixed):
    N, S = 128, 16
    uv, vals = O.gen_uniform(1500, 0, N, seed=4)
    gcf = O.init_gcf(S)
    e = S * (S - 1) // 2 - 1
    r2, i2 = O.compute_gcf_phase(uv, gcf[e, 0], gcf[e, 1], fixed_point=fixed)
    ref = cplx(O.grid_const_tap(uv, vals, N, S, r2, i2))
    g = run(uv, vals, gcf, N, S, planar=strategy == "window", strategy=strategy,
            gcf_mode=migrid.GCF_COMPAT)
    # The phase is a float expression; a 1-ulp difference in sqrtf or in the
    # sin/cos argument reduction moves the tap value by about |phase| * 2^-24.
    # The fixed-point form's dx is about -7*main (the reference feeds the int2
    # fixed point into the formula), so its phase reaches hundreds of radians.
    f32 = np.float32
    p1 = f32(2 * 3.1415926 * gcf[e, 1])
    p2 = f32(float(p1) * gcf[e, 0])
    m = np.floor(uv)
    sub = np.floor(8 * (uv - m))
    base = m * 8 + sub if fixed else uv
    d = (m - base + sub / 8.0).astype(f32)
    phase = np.abs(p1 - p2 * np.sqrt((d * d).sum(1)))
    tol = max(1e-6, 8 * float(phase.max()) * 2.0 ** -24)
    assert O.rel_errors(g, ref)[0] < tol


@pytest.mark.parametrize("strategy", STRATS)
def test_analytic_gcf(strategy):
    N, S = 160, 64
    uv, vals = O.gen_uniform(3000, 0, N, seed=13)
    ref = cplx(O.grid_zz(uv, vals, N, O.config_gcf(S, beta=0.05), S))
    g = run(uv, vals, None, N, S, planar=strategy == "window", strategy=strategy,
            gcf_mode=migrid.GCF_ANALYTIC, beta=0.05)
    assert O.rel_errors(g, ref)[0] < 1e-12


def bf16_round(a):
    b = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("S", [16, 64])
def test_mixed_precision_errors(S):
    N = 256
    uv, vals = O.gen_clustered(30000, N, S, mean_r=40.0, seed=21)
    gcf = O.config_gcf(S)
    ref = cplx(O.grid_zz(uv, vals, N, gcf, S))
    # FP32 products and tile sums, FP64 flush
    g32 = run(uv, vals, gcf, N, S, strategy="tiled", arith=migrid.ARITH_F32)
    assert O.rel_errors(g32, ref)[0] < 1e-5
    # FP32 grid with two-sum compensation
    gc = run(uv, vals, gcf, N, S, out_dtype="f32", grid_lo=True, strategy="tiled",
             arith=migrid.ARITH_F32, vis_per_wave=64)
    assert O.rel_errors(gc, ref)[0] < 1e-5
    # BF16 GCF: within FP32 rounding of the oracle on the bf16-rounded table
    gb = run(uv, vals, gcf, N, S, strategy="tiled", arith=migrid.ARITH_F32_BF16GCF)
    ref_b = cplx(O.grid_zz(uv, vals, N, bf16_round(gcf), S))
    assert O.rel_errors(gb, ref_b)[0] < 1e-5
    assert O.rel_errors(gb, ref)[0] < 2e-2


def test_accumulate_and_tiled_support_limits():
    t = torch()
    N, S = 128, 32
    uv, vals = O.gen_uniform(3000, 0, N, seed=2)
    gcf = O.config_gcf(S)
    ref = cplx(O.grid_zz(uv, vals, N, gcf, S))
    grid = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    ctx = migrid.Context()
    ctx.grid(grid, dev(uv), dev(vals), dev(gcf), N, S, strategy="tiled")
    ctx.grid(grid, dev(uv), dev(vals), dev(gcf), N, S, strategy="gather", accumulate=True)
    t.cuda.synchronize()
    g = grid.cpu().numpy()
    assert O.rel_errors(g[..., 0] + 1j * g[..., 1], 2 * ref)[0] < 1e-12
    with pytest.raises(migrid.MigridError):
        ctx.grid(t.empty((64, 64, 2), dtype=t.float64, device="cuda"), dev(uv[:10]), dev(vals[:10]),
                 dev(O.config_gcf(128)), 64, 128, strategy="tiled")


def test_drop_in_strategies_through_host_api():
    N, S = 200, 32
    uv, vals = O.ref_inputs(5000, N)
    gcf = O.init_gcf(S)
    ref = cplx(O.grid_zz(uv, vals, N, gcf, S))
    for strategy in STRATS:
        out = migrid.gridGPU(uv, vals, gcf, N, S, strategy=strategy)
        planar = strategy == "window"
        g = migrid.unpad(out, N, S, planar=planar)
        if planar:
            g = g[0] + 1j * g[1]
        assert O.rel_errors(g, ref)[0] < 1e-12, strategy


@pytest.mark.parametrize("strategy", ["gather", "window"])
def test_gather_and_window_are_bitwise_reproducible(strategy):
    """Gather and window split dense tiles into work items too, but sum the
    pieces of a split tile in a fixed order instead of with atomics: every
    run gives the same grid bit for bit, within 1e-12 of the oracle, fresh
    or accumulating."""
    t = torch()
    N, S = 200, 32
    uv, vals = O.gen_clustered(40000, N, S, mean_r=6.0, seed=21)  # a few very dense tiles
    gcf = O.init_gcf(S)
    ref = cplx(O.grid_zz(uv, vals, N, gcf, S))
    planar = strategy == "window"
    runs = [run(uv, vals, gcf, N, S, planar=planar, strategy=strategy, vis_per_wave=512)
            for _ in range(3)]
    for g in runs[1:]:
        assert np.array_equal(g, runs[0])
    assert O.rel_errors(runs[0], ref)[0] < 1e-12
    # accumulating into a filled grid
    base = np.random.default_rng(3).normal(size=(2, N, N))
    grid = dev(base if planar else np.moveaxis(base, 0, -1))
    ctx = migrid.Context()
    ctx.grid(grid, dev(uv), dev(vals), dev(gcf), N, S, strategy=strategy, vis_per_wave=512,
             accumulate=True)
    t.cuda.synchronize()
    g = grid.cpu().numpy()
    got = (g[0] + 1j * g[1]) if planar else (g[..., 0] + 1j * g[..., 1])
    assert O.rel_errors(got - (base[0] + 1j * base[1]), ref)[0] < 1e-12


@pytest.mark.parametrize("S,N,strategy", [(2, 50, "gather"), (6, 90, "gather"), (24, 150, "tiled"),
                                          (34, 120, "gather"), (40, 160, "window"), (62, 200, "tiled")])
def test_embedded_supports(S, N, strategy):
    """Supports other than 32 and 64 run on the column kernel embedded in the
    next of them (the table surrounded by zero taps, every tap on its own
    pixel), from the smallest to the largest gap.  FP64 within 1e-12 of
    the oracle, FP32 tiles (tiled strategy) within 1e-5; the grid starts as
    NaN, so a tile left unwritten fails."""
    print(f"[embedded support] S={S} N={N} {strategy}", flush=True)
    uv, vals = O.gen_uniform(3000, -S / 2, N + S / 2, seed=S + 11)
    gcf = O.init_gcf(S)
    ref = cplx(O.grid_zz(uv, vals, N, gcf, S))
    t = torch()
    planar = strategy == "window"
    ariths = [(migrid.ARITH_NATIVE, 1e-12)] + ([(migrid.ARITH_F32, 1e-5)] if strategy == "tiled" else [])
    ctx = migrid.Context()
    for arith, tol in ariths:
        grid = t.full((2, N, N) if planar else (N, N, 2), float("nan"), dtype=t.float64, device="cuda")
        ctx.grid(grid, dev(uv), dev(vals), dev(gcf), N, S, strategy=strategy, arith=arith)
        t.cuda.synchronize()
        g = grid.cpu().numpy()
        g = g[0] + 1j * g[1] if planar else g[..., 0] + 1j * g[..., 1]
        assert O.rel_errors(g, ref)[0] < tol, (arith, S)


@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("S", [32, 24])
def test_sparse_support32_pairs(S, strategy):
    """At S = 32 (and supports embedded in it) the column kernel stages two
    sub-kernels per barrier pair.  A few visibilities in one corner of the
    grid give tiles with one, two or an odd number of non-empty sub-kernels
    (the last pair half empty), items cut down to single sub-kernels, and
    one run where every visibility has the same sub-kernel.  FP64 within
    1e-12 of the oracle, FP32 tiles within 1e-5; the grid starts as NaN."""
    t = torch()
    N = 192
    gcf = O.init_gcf(S)
    planar = strategy == "window"
    ctx = migrid.Context()
    rng = np.random.default_rng(S * 3 + len(strategy))
    cases = []
    for n in (1, 2, 3, 5, 9, 33):
        uv = rng.uniform(40.0, 52.0, size=(n, 2))
        vals = rng.standard_normal((n, 2))
        cases.append((f"n={n}", uv, vals))
    uv = np.floor(rng.uniform(40.0, 140.0, size=(7, 2))) + 0.3125
    cases.append(("one sub-kernel", uv, rng.standard_normal((7, 2))))
    for name, uv, vals in cases:
        print(f"[support-32 pairs] S={S} {strategy} {name}", flush=True)
        ref = cplx(O.grid_zz(uv, vals, N, gcf, S))
        runs = [(migrid.ARITH_NATIVE, 1e-12, {})]
        if strategy == "tiled":
            runs += [(migrid.ARITH_NATIVE, 1e-12, {"vis_per_wave": 1}), (migrid.ARITH_F32, 1e-5, {})]
        for arith, tol, kw in runs:
            grid = t.full((2, N, N) if planar else (N, N, 2), float("nan"), dtype=t.float64,
                          device="cuda")
            ctx.grid(grid, dev(uv), dev(vals), dev(gcf), N, S, strategy=strategy, arith=arith, **kw)
            t.cuda.synchronize()
            g = grid.cpu().numpy()
            g = g[0] + 1j * g[1] if planar else g[..., 0] + 1j * g[..., 1]
            assert O.rel_errors(g, ref)[0] < tol, (name, arith, kw)


def test_job_past_one_plan_is_chunked():
    """A job whose packed records would pass 2^32 bytes (the column kernel's
    byte offsets are 32-bit) runs as consecutive chunks of at most
    2^32 / 32 - 1 visibilities (FP64 records), the later ones added to the
    grid: with GATHER the result is bit for bit the one two calls give, the
    first on the leading chunk into a fresh grid, the second on the rest
    accumulating."""
    t = torch()
    lim = ((1 << 32) - 1) // 32
    n = lim + 1000
    N, S = 256, 8
    print(f"[chunked job] n={n}", flush=True)
    g = t.Generator(device="cuda")
    g.manual_seed(3)
    uv = t.rand((n, 2), dtype=t.float64, device="cuda", generator=g) * (N - 2 * S) + S
    vals = t.randn((n, 2), dtype=t.float64, device="cuda", generator=g)
    gcf = dev(O.init_gcf(S))
    ctx = migrid.Context()
    whole = t.full((N, N, 2), float("nan"), dtype=t.float64, device="cuda")
    ctx.grid(whole, uv, vals, gcf, N, S, strategy="gather", layout=migrid.LAYOUT_INTERLEAVED)
    parts = t.full((N, N, 2), float("nan"), dtype=t.float64, device="cuda")
    ctx.grid(parts, uv[:lim], vals[:lim], gcf, N, S, strategy="gather",
             layout=migrid.LAYOUT_INTERLEAVED)
    ctx.grid(parts, uv[lim:], vals[lim:], gcf, N, S, strategy="gather",
             layout=migrid.LAYOUT_INTERLEAVED, accumulate=True)
    t.cuda.synchronize()
    assert t.isfinite(whole).all()
    assert t.equal(whole, parts)
    del uv, vals
