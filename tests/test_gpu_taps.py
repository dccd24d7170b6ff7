"""GPU parity of the small-support kernel: the TILED strategy at S <= 16 grids
one tap per lane into an LDS tile (k_taps) instead of 16 x 16 wave blocks.

Against the CPU oracle: FP64 within 1e-12 relative max-norm; FP32 tiles
within 1e-5; BF16 GCF tiles within 1e-5 of the oracle on the bf16-rounded
table; float pairs within 2e-5.  Against the owner kernel on the same inputs
(tap_smax = -1 turns the tap kernel off): FP64 within 1e-13.  Dense clusters
with tiny work items force split tiles (atomic flushes onto zeroed tiles);
a batch of small supports replays as a graph, and a batch mixing small and
large supports runs both kernels, each from its own queue.  MIGRID_VERBOSE names the
kernel each launch used, which the tests read to be sure which path ran.
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


def cplx(p):
    return p[0] + 1j * p[1]


def bf16_round(a):
    f = np.ascontiguousarray(a, np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def inputs(N, S, n, dense, seed):
    rng = np.random.default_rng(seed)
    if dense:
        uv = rng.uniform(0, N, size=2) + rng.normal(scale=2.0, size=(n, 2))
    else:
        uv = rng.uniform(-S, N + S, size=(n, 2))
    return uv.astype(np.float32).astype(np.float64), rng.normal(size=(n, 2))


def grid_on(uv, vals, gcf, N, S, out="f64", in_dtype="f64", planar=False, base=None, **kw):
    t = torch()
    dt = t.float64 if out == "f64" else t.float32
    shape = (2, N, N) if planar else (N, N, 2)
    grid = dev(base.astype(np.float64)).to(dt) if base is not None else t.empty(shape, dtype=dt,
                                                                               device="cuda")
    idt = np.float64 if in_dtype == "f64" else np.float32
    ctx = migrid.Context()
    ctx.grid(grid, dev(uv.astype(idt)), dev(vals.astype(idt)), dev(gcf.astype(idt)), N, S,
             layout=migrid.LAYOUT_PLANAR if planar else migrid.LAYOUT_INTERLEAVED,
             accumulate=base is not None, **kw)
    t.cuda.synchronize()
    g = grid.cpu().numpy().astype(np.float64)
    return g[0] + 1j * g[1] if planar else g[..., 0] + 1j * g[..., 1]


def rel(got, want):
    m = np.abs(want).max(initial=0.0)
    if m == 0.0:
        return float(np.abs(got).max(initial=0.0))
    return float(np.abs(got - want).max() / m)


CASES = [(40, 2, 500, False, 16), (100, 4, 3000, True, 32), (129, 8, 4000, False, 64),
         (64, 8, 6000, True, 64), (90, 10, 2000, False, 32), (257, 12, 5000, True, 64),
         (200, 16, 5000, False, 64), (31, 16, 300, False, 16), (5, 6, 50, False, 16)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "N%d_S%d_n%d%s_T%d" % (
    c[0], c[1], c[2], "_dense" if c[3] else "", c[4]))
def test_taps_against_oracle(case, capfd, monkeypatch):
    N, S, n, dense, T = case
    print(f"[taps] case {case}", flush=True)
    monkeypatch.setenv("MIGRID_VERBOSE", "1")
    uv, vals = inputs(N, S, n, dense, seed=N * 100 + S)
    gcf = O.config_gcf(S)
    want = cplx(O.grid_zz(uv, vals, N, gcf, S))
    g = grid_on(uv, vals, gcf, N, S, strategy="tiled", tile=T)
    assert rel(g, want) < 1e-12, "fp64"
    assert "tap launch" in capfd.readouterr().err
    # FP32 tiles, planar, accumulating into a filled grid
    base = np.random.default_rng(S).normal(size=(2, N, N))
    g = grid_on(uv, vals, gcf, N, S, planar=True, base=base, strategy="tiled", tile=T,
                arith=migrid.ARITH_F32) - cplx(base)
    m = max(np.abs(want).max(initial=0.0), 1e-300)
    assert np.abs(g - want).max(initial=0.0) <= 1e-5 * m + 1e-12 * np.abs(base).max()
    # BF16 GCF tiles against the oracle on the rounded table
    wantb = cplx(O.grid_zz(uv, vals, N, bf16_round(gcf), S))
    g = grid_on(uv, vals, gcf, N, S, strategy="tiled", tile=T, arith=migrid.ARITH_F32_BF16GCF)
    assert rel(g, wantb) < 1e-5, "bf16"
    # float pairs: <double2,float2> (float products summed in double), <float2,float2>
    g = grid_on(uv, vals, gcf, N, S, in_dtype="f32", strategy="tiled", tile=T)
    assert rel(g, want) < 2e-5, "zc"
    g = grid_on(uv, vals, gcf, N, S, out="f32", in_dtype="f32", strategy="tiled", tile=T)
    assert rel(g, want) < 2e-5, "cc"
    assert capfd.readouterr().err.count("tap launch") == 4


@pytest.mark.parametrize("S", [2, 8, 12, 16])
def test_taps_match_owner_kernel(S, capfd, monkeypatch):
    N, n = 160, 6000
    print(f"[taps vs owner] S={S}", flush=True)
    monkeypatch.setenv("MIGRID_VERBOSE", "1")
    uv, vals = inputs(N, S, n, True, seed=S)
    gcf = O.config_gcf(S)
    a = grid_on(uv, vals, gcf, N, S, strategy="tiled")
    b = grid_on(uv, vals, gcf, N, S, strategy="tiled", tap_smax=-1)
    err = capfd.readouterr().err
    assert err.count("tap launch") == 1 and err.count("owner launch") == 1
    assert rel(a, b) < 1e-13


def test_tap_smax_threshold(capfd, monkeypatch):
    """tap_smax caps the supports sent to the tap kernel; S = 18 never uses it."""
    monkeypatch.setenv("MIGRID_VERBOSE", "1")
    N = 120
    for S, opt, path in ((12, 8, "owner"), (12, 12, "tap"), (8, 0, "tap"), (18, 0, "owner"),
                         (18, 64, "owner")):
        print(f"[tap_smax] S={S} tap_smax={opt}", flush=True)
        uv, vals = inputs(N, S, 2000, False, seed=S + opt)
        gcf = O.config_gcf(S)
        g = grid_on(uv, vals, gcf, N, S, strategy="tiled", tap_smax=opt)
        assert rel(g, cplx(O.grid_zz(uv, vals, N, gcf, S))) < 1e-12
        err = capfd.readouterr().err
        assert f"{path} launch" in err, (S, opt, err[-300:])


def test_gather_and_window_keep_the_owner_kernel(capfd, monkeypatch):
    monkeypatch.setenv("MIGRID_VERBOSE", "1")
    N, S = 100, 8
    uv, vals = inputs(N, S, 3000, True, seed=1)
    gcf = O.config_gcf(S)
    want = cplx(O.grid_zz(uv, vals, N, gcf, S))
    for strategy in ("gather", "window"):
        planar = strategy == "window"
        a = grid_on(uv, vals, gcf, N, S, planar=planar, strategy=strategy)
        b = grid_on(uv, vals, gcf, N, S, planar=planar, strategy=strategy)
        assert rel(a, want) < 1e-12 and np.array_equal(a, b)
    err = capfd.readouterr().err
    assert "tap launch" not in err and "owner launch" in err


@pytest.mark.parametrize("vpw", [16, 64])
def test_split_tiles(vpw):
    """Tiny work items split the dense tiles over many items, which flush with
    atomics onto tiles the work-item pass zeroed."""
    N, S = 96, 8
    print(f"[taps split] vis_per_wave={vpw}", flush=True)
    uv, vals = inputs(N, S, 20000, True, seed=vpw)
    gcf = O.config_gcf(S)
    want = cplx(O.grid_zz(uv, vals, N, gcf, S))
    for arith, tol in ((migrid.ARITH_NATIVE, 1e-12), (migrid.ARITH_F32, 1e-5)):
        g = grid_on(uv, vals, gcf, N, S, strategy="tiled", vis_per_wave=vpw, arith=arith)
        assert rel(g, want) < tol, arith


def test_small_support_batch_and_graph(capfd, monkeypatch):
    t = torch()
    monkeypatch.setenv("MIGRID_VERBOSE", "1")
    shapes = [(128, 8, 3000), (200, 16, 5000), (64, 2, 800), (150, 12, 0), (300, 4, 4000)]
    jobs, wants = [], []
    for i, (N, S, n) in enumerate(shapes):
        uv, vals = inputs(N, S, n, i % 2 == 1, seed=40 + i)
        gcf = O.config_gcf(S)
        wants.append(cplx(O.grid_zz(uv, vals, N, gcf, S)))
        jobs.append(dict(grid=t.empty((N, N, 2), dtype=t.float64, device="cuda"), uv=dev(uv),
                         vals=dev(vals), gcf=dev(gcf), N=N, S=S))
    ctx = migrid.Context()
    for arith, tol in ((migrid.ARITH_NATIVE, 1e-12), (migrid.ARITH_F32, 1e-5)):
        print(f"[taps batch] arith {arith}", flush=True)
        b = migrid.Batch(ctx, jobs, arith=arith)
        stream = t.cuda.Stream()
        for mode in ("run", "graph", "graph", "graph"):
            for j in jobs:
                j["grid"].fill_(7.0)
            t.cuda.synchronize()
            if mode == "run":
                b.run(stream)
            else:
                b.graph_launch(stream)
            t.cuda.synchronize()
            for j, w in zip(jobs, wants):
                g = j["grid"].cpu().numpy()
                assert rel(g[..., 0] + 1j * g[..., 1], w) < tol, (mode, j["N"], j["S"])
        b.close()
    assert "tap launch" in capfd.readouterr().err


def test_mixed_batch_uses_both_kernels(capfd, monkeypatch):
    """A batch mixing small and large supports: the small ones go to the tap
    kernel's queue, the rest to the owner kernel's; two jobs of different
    supports share one grid (atomic flushes into the caller's zeros)."""
    t = torch()
    monkeypatch.setenv("MIGRID_VERBOSE", "1")
    shapes = [(128, 8, 3000), (200, 40, 4000), (96, 4, 1000), (150, 64, 2000), (180, 16, 5000)]
    jobs, wants = [], []
    for i, (N, S, n) in enumerate(shapes):
        uv, vals = inputs(N, S, n, i % 2 == 0, seed=70 + i)
        gcf = O.config_gcf(S)
        wants.append(cplx(O.grid_zz(uv, vals, N, gcf, S)))
        jobs.append(dict(grid=t.empty((N, N, 2), dtype=t.float64, device="cuda"), uv=dev(uv),
                         vals=dev(vals), gcf=dev(gcf), N=N, S=S))
    N = 140
    shared = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    want_shared = np.zeros((N, N), complex)
    for i, S in enumerate((8, 32)):
        uv, vals = inputs(N, S, 3000, True, seed=90 + i)
        gcf = O.config_gcf(S)
        want_shared += cplx(O.grid_zz(uv, vals, N, gcf, S))
        jobs.append(dict(grid=shared, uv=dev(uv), vals=dev(vals), gcf=dev(gcf), N=N, S=S))
    ctx = migrid.Context()
    for arith, tol in ((migrid.ARITH_NATIVE, 1e-12), (migrid.ARITH_F32, 1e-5)):
        print(f"[mixed batch] arith {arith}", flush=True)
        b = migrid.Batch(ctx, jobs, arith=arith)
        stream = t.cuda.Stream()
        for mode in ("run", "graph", "graph"):
            for j in jobs[:len(shapes)]:
                j["grid"].fill_(3.0)
            shared.zero_()  # jobs sharing a grid add into the caller's grid
            t.cuda.synchronize()
            if mode == "run":
                b.run(stream)
            else:
                b.graph_launch(stream)
            t.cuda.synchronize()
            for j, w in zip(jobs, wants):
                g = j["grid"].cpu().numpy()
                assert rel(g[..., 0] + 1j * g[..., 1], w) < tol, (mode, j["N"], j["S"])
            g = shared.cpu().numpy()
            assert rel(g[..., 0] + 1j * g[..., 1], want_shared) < tol, (mode, "shared")
        b.close()
    err = capfd.readouterr().err
    assert "tap launch" in err and "owner launch" in err


def sweep_cases(n=24, seed=20261204):
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(n):
        N = int(rng.choice([3, 7, 16, 33, 64, 129, 300, int(rng.integers(4, 400))]))
        S = int(2 * rng.integers(1, 9))
        npts = int(rng.choice([0, 1, 17, 1000, 5000]))
        dense = bool(rng.random() < 0.4) and npts > 0
        tile = int(rng.choice([16, 32, 64]))
        arith = int(rng.choice([migrid.ARITH_NATIVE, migrid.ARITH_F32, migrid.ARITH_F32_BF16GCF]))
        planar = bool(rng.random() < 0.5)
        accumulate = bool(rng.random() < 0.3)
        cases.append((i, N, S, npts, dense, tile, arith, planar, accumulate))
    return cases


@pytest.mark.parametrize("case", sweep_cases(), ids=lambda c: "i%d_N%d_S%d_n%d%s_T%d_a%d%s%s" % (
    c[0], c[1], c[2], c[3], "_dense" if c[4] else "", c[5], c[6], "_planar" if c[7] else "",
    "_acc" if c[8] else ""))
def test_taps_random_sweep(case):
    """Seeded shapes through the tap kernel: grid edges below the support,
    empty and single-visibility jobs, dense clusters, every tile edge, the
    three arithmetic forms, both layouts, fresh and accumulating grids."""
    i, N, S, npts, dense, tile, arith, planar, accumulate = case
    print(f"[taps sweep] {case}", flush=True)
    uv, vals = inputs(N, S, npts, dense, seed=5000 + i)
    gcf = O.config_gcf(S)
    table = bf16_round(gcf) if arith == migrid.ARITH_F32_BF16GCF else gcf
    want = cplx(O.grid_zz(uv, vals, N, table, S))
    base = np.random.default_rng(i).normal(size=(2, N, N)) if accumulate else None
    if base is not None and not planar:
        base = np.moveaxis(base, 0, -1)
    g = grid_on(uv, vals, gcf, N, S, planar=planar, base=base, strategy="tiled", tile=tile,
                arith=arith)
    if base is not None:
        g = g - (cplx(base) if planar else base[..., 0] + 1j * base[..., 1])
    tol = 1e-12 if arith == migrid.ARITH_NATIVE else 1e-5
    m = np.abs(want).max(initial=0.0)
    if m == 0.0 and base is None:
        assert np.abs(g).max(initial=0.0) == 0.0
    else:
        scale = max(m, 1e-300)
        extra = 1e-12 * np.abs(base).max() if base is not None else 0.0
        assert np.abs(g - want).max(initial=0.0) <= tol * scale + extra, f"rel {rel(g, want):.3e}"


@pytest.mark.parametrize("S", [8, 16])
def test_taps_coverage_bit_exact(S, capfd, monkeypatch):
    """The DEBUG1 coverage map (every tap adds (1, n)) through the tap kernel:
    sums of integers in double, equal to the oracle's bit for bit."""
    monkeypatch.setenv("MIGRID_VERBOSE", "1")
    N = 120
    uv, vals = O.gen_uniform(3000, -S, N + S, seed=S + 3)
    ref = O.grid_zz(uv, vals, N, O.config_gcf(S), S, debug1=True)
    g = grid_on(uv, vals, O.config_gcf(S), N, S, strategy="tiled", debug_coverage=True)
    np.testing.assert_array_equal(g.real, ref[0])
    np.testing.assert_array_equal(g.imag, ref[1])
    assert "tap launch" in capfd.readouterr().err


@pytest.mark.parametrize("S", [4, 8, 16])
def test_taps_compute_gcf_matches_owner(S, capfd, monkeypatch):
    """COMPUTE_GCF (the constant-tap form, values carrying the phase) on the
    tap kernel against the owner kernel on the same inputs."""
    monkeypatch.setenv("MIGRID_VERBOSE", "1")
    N = 140
    uv, vals = inputs(N, S, 4000, True, seed=S + 11)
    gcf = O.init_gcf(S)
    a = grid_on(uv, vals, gcf, N, S, strategy="tiled", gcf_mode=migrid.GCF_COMPAT)
    b = grid_on(uv, vals, gcf, N, S, strategy="tiled", gcf_mode=migrid.GCF_COMPAT, tap_smax=-1)
    err = capfd.readouterr().err
    assert err.count("tap launch") == 1 and err.count("owner launch") == 1
    assert rel(a, b) < 1e-13


@pytest.mark.parametrize("S", [8, 16])
@pytest.mark.parametrize("arith,tol", [(migrid.ARITH_NATIVE, 1e-13), (migrid.ARITH_F32, 1e-5)])
def test_taps_analytic_gcf_matches_owner(S, arith, tol, capfd, monkeypatch):
    """The analytic GCF (Gaussian x chirp, evaluated on the device) on the tap
    kernel against the owner kernel, which stages the same factors."""
    monkeypatch.setenv("MIGRID_VERBOSE", "1")
    t = torch()
    N = 130
    uv, vals = inputs(N, S, 3000, False, seed=S + 23)
    grids = []
    for tap_smax in (0, -1):
        g = t.empty((N, N, 2), dtype=t.float64, device="cuda")
        migrid.Context().grid(g, dev(uv), dev(vals), None, N, S, strategy="tiled",
                              gcf_mode=migrid.GCF_ANALYTIC, sigma=S / 6.0, beta=0.05,
                              arith=arith, tap_smax=tap_smax)
        t.cuda.synchronize()
        h = g.cpu().numpy()
        grids.append(h[..., 0] + 1j * h[..., 1])
    err = capfd.readouterr().err
    assert "tap launch" in err and "owner launch" in err
    assert rel(grids[0], grids[1]) < tol
