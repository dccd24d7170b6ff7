"""Pins the CPU oracle (oracle/gridcpu.c) without trusting it.

The reference ships no golden vectors, and its gridCPU cannot be compiled
here without a stand-in for cuda.h (grid.cu:35), so every check below comes
from something the oracle's own code does not share: closed forms (impulse
response = the addressed sub-kernel, grid.cu:115-132), an independent numpy
restatement by window slicing, integer coverage counts (the reference's
DEBUG1 mode, grid.cu:127-129), linearity, brute force over tiny inputs and
direct calls into glibc for the reference's input stream (grid.cu:444-452).
"""
import ctypes
import math

import numpy as np
import pytest

import oracle as O


def numpy_grid(uv, vals, N, gcf, S):
    """Independent restatement: place each visibility's whole S x S window
    by slicing, visibilities in ascending order (so per-pixel sums run in the
    same order as the reference loop and the result is bit-identical)."""
    h = S // 2
    re = np.zeros((N, N))
    im = np.zeros((N, N))
    G = np.asarray(gcf, np.float64).reshape(64, S, S, 2)
    for (u, v), (r1, i1) in zip(np.asarray(uv), np.asarray(vals)):
        mx, my = math.floor(u), math.floor(v)
        sx, sy = math.floor(8 * (u - mx)), math.floor(8 * (v - my))
        g = G[8 * sy + sx]
        # window rows y = my-h+1 .. my+h, cols x = mx-h+1 .. mx+h
        y0, x0 = my - h + 1, mx - h + 1
        ys = slice(max(y0, 0), min(y0 + S, N))
        xs = slice(max(x0, 0), min(x0 + S, N))
        if ys.start >= ys.stop or xs.start >= xs.stop:
            continue
        gr = g[ys.start - y0: ys.stop - y0, xs.start - x0: xs.stop - x0]
        re[ys, xs] += r1 * gr[..., 0] - i1 * gr[..., 1]
        im[ys, xs] += r1 * gr[..., 1] + gr[..., 0] * i1
    return np.stack([re, im])


@pytest.mark.parametrize("S", [8, 16, 32])
def test_impulse_is_the_addressed_subkernel(S):
    N = 4 * S
    G = O.init_gcf(S).reshape(64, S, S, 2)  # asymmetric: catches transposes
    h = S // 2
    for sx, sy, mx, my in [(0, 0, 2 * S, 2 * S), (3, 5, 2 * S + 1, 2 * S - 3), (7, 1, S, 3 * S - 1)]:
        uv = np.array([[mx + sx / 8.0, my + sy / 8.0]])
        vals = np.array([[1.0, 0.0]])
        out = O.grid_zz(uv, vals, N, G.reshape(-1, 2), S)
        win = out[:, my - h + 1: my + h + 1, mx - h + 1: mx + h + 1]
        k = 8 * sy + sx
        np.testing.assert_array_equal(win[0], G[k, :, :, 0])
        np.testing.assert_array_equal(win[1], G[k, :, :, 1])
        # nothing outside the window
        mask = np.ones((N, N), bool)
        mask[my - h + 1: my + h + 1, mx - h + 1: mx + h + 1] = False
        assert not out[:, mask].any()
        # imaginary unit value rotates the kernel: (0+1i)*g = -gi + i gr
        out_i = O.grid_zz(uv, np.array([[0.0, 1.0]]), N, G.reshape(-1, 2), S)
        wi = out_i[:, my - h + 1: my + h + 1, mx - h + 1: mx + h + 1]
        np.testing.assert_array_equal(wi[0], -G[k, :, :, 1])
        np.testing.assert_array_equal(wi[1], G[k, :, :, 0])


@pytest.mark.parametrize("S,N,npts", [(8, 48, 300), (16, 64, 200), (32, 96, 120)])
def test_matches_independent_slicing_restatement_bitwise(S, N, npts):
    # uv spans beyond the grid on every side, so edge clipping is exercised
    uv, vals = O.gen_uniform(npts, -S / 2.0, N + S / 2.0, seed=11 + S)
    gcf = O.config_gcf(S)
    got = O.grid_zz(uv, vals, N, gcf, S)
    ref = numpy_grid(uv, vals, N, gcf, S)
    np.testing.assert_array_equal(got, ref)
    gi = O.init_gcf(S)
    np.testing.assert_array_equal(O.grid_zz(uv, vals, N, gi, S), numpy_grid(uv, vals, N, gi, S))


def test_coverage_counts_debug1():
    S, N, npts = 16, 80, 500
    uv, vals = O.gen_uniform(npts, -10, N + 10, seed=5)
    out = O.grid_zz(uv, vals, N, O.config_gcf(S), S, debug1=True)
    # independent count: pixel (x,y) is covered by vis n iff
    # floor(u)-S/2+1 <= x <= floor(u)+S/2 (and the same in y)
    m = np.floor(uv).astype(np.int64)
    xs = np.arange(N)
    cx = (xs[None, :] >= m[:, 0:1] - S // 2 + 1) & (xs[None, :] <= m[:, 0:1] + S // 2)
    cy = (xs[None, :] >= m[:, 1:2] - S // 2 + 1) & (xs[None, :] <= m[:, 1:2] + S // 2)
    cnt = np.einsum("ny,nx->yx", cy.astype(np.int64), cx.astype(np.int64))
    idx = np.einsum("ny,nx->yx", cy * np.arange(npts)[:, None], cx.astype(np.int64))
    np.testing.assert_array_equal(out[0], cnt.astype(np.float64))
    np.testing.assert_array_equal(out[1], idx.astype(np.float64))


def test_linearity_and_pz_accumulation():
    S, N = 16, 64
    uv, v1 = O.gen_uniform(200, 0, N, seed=1)
    _, v2 = O.gen_uniform(200, 0, N, seed=2)
    gcf = O.config_gcf(S)
    g1 = O.grid_zz(uv, v1, N, gcf, S)
    g2 = O.grid_zz(uv, v2, N, gcf, S)
    g12 = O.grid_zz(uv, 2.5 * v1 + v2, N, gcf, S)
    assert O.rel_errors(g12, 2.5 * g1 + g2)[0] < 1e-13
    # gridCPU_pz adds into what is there (grid.cu:150 does not zero)
    acc = g1.copy()
    O.grid_zz(uv, v2, N, gcf, S, out=acc)
    assert O.rel_errors(acc, g1 + g2)[0] < 1e-14


def test_row_bands_and_openmp_are_bit_identical():
    S, N = 32, 256
    uv, vals = O.gen_clustered(3000, N, S, mean_r=40.0, seed=9)
    gcf = O.config_gcf(S)
    ref = O.grid_zz(uv, vals, N, gcf, S)
    got, used = O.grid_zz_omp(uv, vals, N, gcf, S, 4)
    assert used >= 1
    np.testing.assert_array_equal(got, ref)
    part = O.grid_zz_rows(uv, vals, N, gcf, S, 0, 100)
    O.grid_zz_rows(uv, vals, N, gcf, S, 100, N, out=part)
    np.testing.assert_array_equal(part, ref)
    # windows crossing every edge of a grid that is not a multiple of the band,
    # for both GCF tables and the constant-tap (COMPUTE_GCF) form
    for S, N in ((16, 77), (64, 50), (2, 9), (64, 300)):
        uv, vals = O.gen_uniform(2000, -S, N + S, seed=S + N)
        for gcf in (O.config_gcf(S), O.init_gcf(S)):
            got, _ = O.grid_zz_omp(uv, vals, N, gcf, S, 3)
            np.testing.assert_array_equal(got, O.grid_zz(uv, vals, N, gcf, S))
        r2, i2 = O.compute_gcf_phase(uv, 0.3, 0.2)
        got, _ = O.grid_const_tap_omp(uv, vals, N, S, r2, i2, 3)
        np.testing.assert_array_equal(got, O.grid_const_tap(uv, vals, N, S, r2, i2))


def test_floorf_emulation_K4():
    # float-representable uv: floorf path == exact path
    uv, _ = O.gen_uniform(1000, 0, 2048, seed=3)
    assert O.floorf_divergent(uv) == 0
    # a double just below an integer rounds up in float: floorf((float)u) = m+1,
    # so 8*(u - (m+1)) < 0 and the reference's sub index becomes -1
    u = 100.0 - 1e-9
    bad = np.array([[u, 50.25]])
    assert O.floorf_divergent(bad) == 1
    # just below a sub-pixel boundary: the float rounding of 8*frac gives 8
    u2 = 7.0 + (1.0 - 2.0 ** -30)
    assert O.floorf_divergent(np.array([[3.5, u2]])) == 1
    # the emulated gridder skips a visibility whose emulated k is outside [0,64)
    S, N = 8, 32
    g_exact = O.grid_zz(np.array([[16.0 - 1e-9, 16.5]]), np.array([[1.0, 0.0]]), N, O.config_gcf(S), S)
    g_emul = O.grid_zz(np.array([[16.0 - 1e-9, 16.5]]), np.array([[1.0, 0.0]]), N, O.config_gcf(S), S,
                       floorf_emul=True)
    assert np.abs(g_exact).sum() > 0
    # sub_y = 4 (exact), sub_x = -1 (emulated) -> k = 31 instead of 39: a
    # different sub-kernel, shifted main pixel unchanged
    assert not np.array_equal(g_exact, g_emul)


def test_init_gcf_closed_form():
    S = 16
    g = O.init_gcf(S)
    for sub_x, sub_y, x, y in [(0, 0, 0, 0), (3, 5, 7, 2), (7, 7, 15, 15), (1, 6, 9, 0)]:
        tmp = math.sin(6.28 * x / S / 8) * math.exp(-(1.0 * x * x + 1.0 * y * y * sub_y) / S / S / 2)
        i = S * S * (sub_x + sub_y * 8) + x + y * S
        assert g[i, 0] == tmp * math.sin(1.0 * x * sub_x / (y + 1)) + 0.4
        assert g[i, 1] == tmp * math.cos(1.0 * x * sub_x / (y + 1)) - 0.2


@pytest.mark.parametrize("S", [8, 64])
def test_config_gcf_properties(S):
    g = O.config_gcf(S)
    G = (g[:, 0] + 1j * g[:, 1]).reshape(64, S, S)
    h = S // 2
    # tap (r,c) = (h-1,h-1) of sub-kernel 0 sits on the visibility: value 1
    assert G[0, h - 1, h - 1] == 1.0
    # separable: every sub-kernel is rank one
    for k in (0, 9, 63):
        s = np.linalg.svd(G[k], compute_uv=False)
        assert s[1] / s[0] < 1e-13
    # sub-kernel (sx,sy) is the transpose of (sy,sx)
    np.testing.assert_allclose(G[8 * 3 + 5], G[8 * 5 + 3].T, rtol=0, atol=1e-15)
    # modulus is the Gaussian of the distance
    sx, sy, r, c = 2, 6, 5, S - 3
    dx, dy = (c - h + 1) - sx / 8, (r - h + 1) - sy / 8
    assert abs(abs(G[8 * sy + sx, r, c]) - math.exp(-(dx * dx + dy * dy) / (2 * (S / 6) ** 2))) < 1e-15


def test_reference_input_stream_is_glibc_rand():
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    uv, vals = O.ref_inputs(50, 2048)
    libc.srand(1541617)
    RAND_MAX = 2147483647
    f32 = np.float32
    for n in range(50):
        x = f32(f32(libc.rand()) / f32(RAND_MAX)) * f32(2048)
        y = f32(f32(libc.rand()) / f32(RAND_MAX)) * f32(2048)
        vr = f32(libc.rand()) / f32(RAND_MAX)
        vi = f32(libc.rand()) / f32(RAND_MAX)
        assert uv[n, 0] == float(x) and uv[n, 1] == float(y)
        assert vals[n, 0] == float(vr) and vals[n, 1] == float(vi)


def test_generators():
    N, S = 1024, 64
    uv, vals = O.gen_clustered(20000, N, S, mean_r=100.0, seed=4)
    m = np.floor(uv)
    assert (m - S // 2 + 1 >= 0).all() and (m + S // 2 <= N - 1).all()
    assert (uv.astype(np.float32).astype(np.float64) == uv).all()
    assert abs(vals.mean()) < 0.02 and abs(vals.std() - 1) < 0.02
    r = np.hypot(uv[:, 0] - N / 2, uv[:, 1] - N / 2)
    assert 90 < r.mean() < 110
    uv2, _ = O.gen_clustered(20000, N, S, mean_r=100.0, seed=4)
    np.testing.assert_array_equal(uv, uv2)
    uvu, _ = O.gen_uniform(5000, 64, 1984, seed=7)
    assert uvu.min() >= 64 and uvu.max() < 1984


def test_vis2ints_and_set_bookmarks():
    N, S = 256, 32
    uv, _ = O.gen_uniform(3000, 0, N, seed=8)
    ints = O.vis2ints(uv)
    np.testing.assert_array_equal(ints[:, 0], np.floor(uv[:, 0]).astype(int) * 8 + np.floor(8 * (uv[:, 0] - np.floor(uv[:, 0]))).astype(int))
    bs = S // 2
    bg = (N + bs - 1) // bs
    cell = (ints[:, 1] // 8 // bs) * bg + ints[:, 0] // 8 // bs
    order = np.argsort(cell, kind="stable")
    ints_s = ints[order]
    bm = O.set_bookmarks(ints_s, bs, bg)
    np.testing.assert_array_equal(bm, np.searchsorted(cell[order], np.arange(bg * bg + 1), side="left"))


@pytest.mark.parametrize("N,S,H", [(128, 16, 16), (200, 32, 32), (256, 64, 32), (64, 8, 8)])
def test_bin_sort(N, S, H):
    uv, _ = O.gen_uniform(4000, -S, N + S, seed=N + S)
    perm, offs = O.bin_sort(uv, N, S, H)
    nb = O.n_bins(N, S, H)
    # independent key: q = floor(u)+S/2 in [0, N+S-2]
    m = np.floor(uv).astype(np.int64)
    sub = np.floor(8 * (uv - m)).astype(np.int64)
    q = m + S // 2
    ok = (q >= 0).all(1) & (q <= N + S - 2).all(1)
    key = (8 * sub[:, 1] + sub[:, 0]) * nb * nb + (q[:, 1] // H) * nb + q[:, 0] // H
    idx = np.nonzero(ok)[0]
    exp_perm = idx[np.argsort(key[idx], kind="stable")]
    np.testing.assert_array_equal(perm, exp_perm)
    np.testing.assert_array_equal(offs, np.searchsorted(np.sort(key[idx]), np.arange(64 * nb * nb + 1)))
    # every dropped visibility really touches no pixel
    for n in np.nonzero(~ok)[0][:50]:
        assert np.abs(numpy_grid(uv[n:n + 1], np.array([[1.0, 1.0]]), N, O.config_gcf(S), S)).sum() == 0


def test_window_pixel_brute_force():
    S, N = 16, 100
    for m in range(-5, N + 5):
        for t in range(S):
            p = O.window_pixel(m, t, S, N)
            cands = [x for x in range(m - S // 2 + 1, m + S // 2 + 1) if x % S == t]
            if m + S // 2 < t:
                assert p == -1
            else:
                assert len(cands) == 1
                assert p == (cands[0] if cands[0] < N else -1)


def _phase_p(w, T):
    """p1 = (float)(2 * 3.1415926 * w) and p2 = (float)(p1 * T): the host-side
    constants of COMPUTE_GCF (grid_gpu.cu:103-108)."""
    p1 = float(np.float32(2 * 3.1415926 * w))
    p2 = float(np.float32(p1 * T))
    return p1, p2


def _assert_tap(r2, i2, phi, err_phi):
    """(r2, i2) = (sin, cos) of an exact phase phi, to within the float
    rounding of the phase (err_phi) and of sinf / cosf (2 ulp of 1)."""
    tol = err_phi + 2.0 ** -22
    np.testing.assert_allclose(r2.astype(np.float64), np.sin(phi), rtol=0, atol=tol)
    np.testing.assert_allclose(i2.astype(np.float64), np.cos(phi), rtol=0, atol=tol)


@pytest.mark.parametrize("form", [0, 1])
def test_compute_gcf_phase_w0_and_T0(form):
    """w = 0 makes p1 = p2 = 0, so every tap is (sin 0, cos 0) = (0, 1) exactly;
    T = 0 makes p2 = 0, so the phase is p1 for every visibility whatever its
    offset: (r2, i2) = (sinf(p1), cosf(p1)) (sin first, grid_gpu.cu:146-147)."""
    uv, _ = O.gen_uniform(200, 0, 500, seed=31 + form)
    r2, i2 = O.compute_gcf_phase(uv, 0.7, 0.0, fixed_point=bool(form))
    np.testing.assert_array_equal(r2, np.zeros(200, np.float32))
    np.testing.assert_array_equal(i2, np.ones(200, np.float32))
    w = 0.31
    p1, _ = _phase_p(w, 0.0)
    r2, i2 = O.compute_gcf_phase(uv, 0.0, w, fixed_point=bool(form))
    assert (r2 == r2[0]).all() and (i2 == i2[0]).all()
    _assert_tap(r2, i2, np.full(200, p1), 0.0)
    assert r2[0] != i2[0]  # (sin, cos), not (cos, sin): p1 is not pi/4


def test_compute_gcf_phase_form0_offsets():
    """Scatter form: dx = (float)(main - u + sub/8) is minus the position's
    remainder below its eighth.  uv on exact eighths: dx = dy = 0, phase = p1
    for any T.  An offset of 1/16 in u alone: |d| = 1/16, phase = p1 - p2/16."""
    T, w = 0.7, 0.31
    p1, p2 = _phase_p(w, T)
    m = np.array([3, 17, 250, 1000])
    s = np.array([0, 3, 5, 7])
    uv = np.stack([m + s / 8.0, m[::-1] + s[::-1] / 8.0], 1)
    r2, i2 = O.compute_gcf_phase(uv, T, w)
    _assert_tap(r2, i2, np.full(4, p1), 0.0)
    uv2 = uv.copy()
    uv2[:, 0] += 1.0 / 16.0
    r2, i2 = O.compute_gcf_phase(uv2, T, w)
    phi = p1 - p2 / 16.0
    _assert_tap(r2, i2, np.full(4, phi), 2.0 ** -22 * 4)
    # the fixed-point form ignores the 1/16 (the fixed point keeps eighths)
    r2f, i2f = O.compute_gcf_phase(uv2, T, w, fixed_point=True)
    r2e, i2e = O.compute_gcf_phase(uv, T, w, fixed_point=True)
    np.testing.assert_array_equal(r2f, r2e)
    np.testing.assert_array_equal(i2f, i2e)


def test_compute_gcf_phase_form1_fixed_point_distance():
    """Gather / moving-window form: u is replaced by the int2 fixed point
    X = 8 main + sub (grid_gpu.cu:456-459, :600-603), so dx = main - X + sub/8
    = -7 X / 8 exactly and |d| = 7/8 sqrt(X^2 + Y^2).  main = sub = 0 gives
    phase p1; (X, Y) on 3-4-5 triangles give exact distances."""
    T, w = 0.7, 0.31
    p1, p2 = _phase_p(w, T)
    # (u, v) -> (X, Y): exact eighths so main and sub are as written
    cases = [((0.0, 0.0), 0.0),                 # X = Y = 0
             ((0.05, 0.1), 0.0),                # still main = sub = 0 (below 1/8)
             ((3 / 8.0, 4 / 8.0), 35 / 8.0),    # X = 3, Y = 4
             ((3.0, 4.0), 35.0),                # X = 24, Y = 32
             ((5.125, 0.0), 7 * 41 / 8.0),      # X = 41, Y = 0
             ((7.5, 10.0), 7 * 100 / 8.0)]      # X = 60, Y = 80: 7/8 * 100
    uv = np.array([c[0] for c in cases])
    dist = np.array([c[1] for c in cases])
    r2, i2 = O.compute_gcf_phase(uv, T, w, fixed_point=True)
    phi = p1 - p2 * dist
    # float rounding of p2 * |d| and of the difference
    err = 2.0 ** -23 * (np.abs(p2 * dist) + np.abs(phi)) * 2
    for n in range(len(cases)):
        _assert_tap(r2[n:n + 1], i2[n:n + 1], phi[n:n + 1], float(err[n]))
    # the scatter form sees the exact eighths at distance 0 (all but (0.05, 0.1))
    ex = np.array([0, 2, 3, 4, 5])
    r0, i0 = O.compute_gcf_phase(uv[ex], T, w)
    _assert_tap(r0, i0, np.full(len(ex), p1), 0.0)


def test_compute_gcf_phase_grid_is_value_times_constant_tap():
    uv, vals = O.gen_uniform(64, 0, 128, seed=12)
    r2, i2 = O.compute_gcf_phase(uv, 0.7, 0.31)
    S, N = 8, 128
    g = O.grid_const_tap(uv, vals, N, S, r2, i2)
    # independent: constant tap value = impulse with an all-ones table times value
    tbl = np.zeros((64 * S * S, 2))
    tbl[:, 0] = 1.0
    v = (vals[:, 0] + 1j * vals[:, 1]) * (r2.astype(np.float64) + 1j * i2.astype(np.float64))
    ref = numpy_grid(uv, np.stack([v.real, v.imag], 1), N, tbl, S)
    assert O.rel_errors(g, ref)[0] < 1e-14
    got, used = O.grid_const_tap_omp(uv, vals, N, S, r2, i2, 3)
    assert used >= 1
    np.testing.assert_array_equal(got, g)


def test_float_pairs_close_to_double():
    S, N = 16, 96
    uv, vals = O.gen_uniform(500, 0, N, seed=21)
    gcf = O.config_gcf(S)
    ref = O.grid_zz(uv, vals, N, gcf, S)
    zc = O.grid_zc(uv, vals, N, gcf, S)
    cc = O.grid_cc(uv, vals, N, gcf, S)
    assert O.rel_errors(zc, ref)[0] < 1e-6
    assert O.rel_errors(cc, ref)[0] < 1e-5
    # zc sums float products in double: equal to a numpy restatement of that
    G = gcf.astype(np.float32).reshape(64, S, S, 2)
    assert zc.dtype == np.float64 and cc.dtype == np.float32
    assert G.shape[0] == 64


def test_uvw_to_pixels_closed_forms():
    c = 299792458.0
    N, cell = 1024, 2.0 ** -20
    # f = c makes u_lambda = u; one uv cell is 1 / (N cell) = 1024 lambda -> one pixel
    uvw = np.array([[0.0, 0.0, 0.0], [1024.0, -2048.0, 5.0], [-512.0, 3072.0, -7.0]])
    uv, _ = O.uvw_to_pixels(uvw, np.array([c]), N, cell)
    np.testing.assert_array_equal(uv, [[512, 512], [513, 510], [511.5, 515]])
    # visibility n = row * nchan + chan; the offset from the centre scales with f
    f = np.array([1.0e8, 2.0e8, 3.0e8])
    uv, _ = O.uvw_to_pixels(uvw, f, N, cell)
    d = (uv - N / 2).reshape(3, 3, 2)
    # (the offset is read back as uv - N/2, so compare to the pixel's rounding)
    np.testing.assert_allclose(d[1, 1], 2 * d[1, 0], rtol=1e-11)
    np.testing.assert_allclose(d[2, 2], 3 * d[2, 0], rtol=1e-11)
    assert d[1, 0, 0] > 0 and d[1, 0, 1] < 0  # u -> x, v -> y, signs kept
    # 1024 m at 100 MHz is 1024e8/c lambda = 1e8/c uv cells of 1024 lambda
    np.testing.assert_allclose(d[1, 0, 0], 1e8 / c, rtol=1e-11)


def test_uvw_to_pixels_w_planes():
    c = 299792458.0
    w = np.array([0.0, 24.999, 25.0, 49.9, 50.0, -74.0, 99.9, 100.0, 1e6])
    uvw = np.stack([np.zeros_like(w), np.zeros_like(w), w], 1)
    _, g = O.uvw_to_pixels(uvw, np.array([c]), 256, 1e-3, nplanes=4, w_max=100.0)
    np.testing.assert_array_equal(g, [0, 0, 1, 1, 2, 2, 3, 3, 3])


@pytest.mark.parametrize("S,N", [(8, 40), (16, 64), (32, 96)])
def test_degrid_is_the_adjoint_of_gridcpu(S, N):
    """<grid(v), g> = <v, degrid(g)> for random v, g (complex inner products):
    a sign, conjugation, index or transpose error in either side breaks it."""
    rng = np.random.default_rng(S)
    uv, v = O.gen_uniform(300, -S / 2, N + S / 2, seed=S + 1)
    gcf = rng.normal(size=(64 * S * S, 2))
    g = rng.normal(size=(2, N, N))
    Av = O.grid_zz(uv, v, N, gcf, S)
    lhs = np.vdot(Av[0] + 1j * Av[1], g[0] + 1j * g[1])
    Ag = O.degrid_zz(uv, g, N, gcf, S)
    rhs = np.vdot(v[:, 0] + 1j * v[:, 1], Ag[:, 0] + 1j * Ag[:, 1])
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


def test_degrid_single_pixel_grid():
    """A grid with one unit pixel p returns, for each visibility, conj of the
    tap of its sub-kernel that lands on p (zero when none does)."""
    S, N = 8, 32
    gcf = O.init_gcf(S)
    G = gcf.reshape(64, S, S, 2)
    g = np.zeros((2, N, N))
    g[0, 13, 17] = 1.0
    uv = np.array([[17.3, 13.8], [20.0, 10.5], [5.0, 5.0], [16.99, 13.0]])
    got = O.degrid_zz(uv, g, N, gcf, S)
    for n, (u, v) in enumerate(uv):
        mx, my = int(np.floor(u)), int(np.floor(v))
        k = 8 * int(np.floor(8 * (v - my))) + int(np.floor(8 * (u - mx)))
        c, r = 17 - (mx - S // 2 + 1), 13 - (my - S // 2 + 1)
        want = G[k, r, c, 0] - 1j * G[k, r, c, 1] if (0 <= c < S and 0 <= r < S) else 0
        assert got[n, 0] + 1j * got[n, 1] == want


@pytest.mark.parametrize("N", [2, 6, 8])
def test_centred_transforms_match_the_definition(N):
    rng = np.random.default_rng(N)
    g = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    assert np.allclose(O.grid_to_image(g), O.centred_dft_direct(g, +1), rtol=0, atol=1e-12)
    assert np.allclose(O.image_to_grid(g), O.centred_dft_direct(g, -1), rtol=0, atol=1e-12)
    t = 0.5 + rng.random(N)
    want = O.centred_dft_direct(g, +1) * 2.5 / np.outer(t, t)
    assert np.allclose(O.grid_to_image(g, 2.5, t), want, rtol=0, atol=1e-11)


def test_centred_transforms_are_adjoint():
    rng = np.random.default_rng(3)
    N = 16
    g = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    h = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    t = 0.5 + rng.random(N)
    lhs = np.vdot(h, O.grid_to_image(g, 0.3, t))
    rhs = np.vdot(O.image_to_grid(h, 0.3, t), g)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("N", [8, 10, 64])
def test_checkerboard_centring_for_even_sizes(N):
    """The device's centring: s . FFT(s . g) with s = (-1)^(x+y) equals the
    shifted transform for even N, and not for odd N."""
    rng = np.random.default_rng(N)
    g = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    s = (-1.0) ** np.add.outer(np.arange(N), np.arange(N))
    assert np.allclose(s * np.fft.ifft2(s * g) * N * N, O.grid_to_image(g), rtol=0, atol=1e-10)
    assert np.allclose(s * np.fft.fft2(s * g), O.image_to_grid(g), rtol=0, atol=1e-10)
    M = N + 1
    h = rng.normal(size=(M, M)) + 0j
    s = (-1.0) ** np.add.outer(np.arange(M), np.arange(M))
    assert not np.allclose(s * np.fft.ifft2(s * h) * M * M, O.centred_dft_direct(h, +1))
