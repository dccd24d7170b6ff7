"""GPU parity of the plain scatter strategy and the bin sort against the oracle.

Each GPU call goes through the C ABI (libmigrid.so via ctypes).  Tolerances:
FP64 grids within 1e-12 relative max-norm of the oracle (atomic ordering
rules out bit-exactness); integer work (coverage counts, bin permutation,
CSR offsets) bit-exact; float pairs within float rounding (stated per test).
"""
import numpy as np
import pytest

import migrid
import oracle as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def dev(a, dtype=None):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


@pytest.mark.parametrize("S,N,npts", [(8, 64, 3000), (16, 96, 2000), (32, 256, 20000), (64, 256, 3000),
                                      (6, 40, 500), (128, 300, 200)])
def test_scatter_zz_matches_gridcpu(S, N, npts):
    uv, vals = O.gen_uniform(npts, -S / 2, N + S / 2, seed=100 + S)  # edges clipped
    gcf = O.init_gcf(S)
    ref = O.grid_zz(uv, vals, N, gcf, S)
    out = migrid.gridGPU(uv, vals, gcf, N, S, pair="zz")
    g = migrid.unpad(out, N, S)
    assert O.rel_errors(g, ref)[0] < 1e-12
    # padding untouched (zeroed, never written)
    base = migrid.grid_base(N, S)
    assert not out[:base].any() and not out[base + N * N:].any()


def test_scatter_reference_input_stream():
    # the reference's own default inputs (glibc rand, grid.cu:444-452) at a
    # reduced count; u can equal N exactly, so edge clipping is exercised
    N, S = 2048, 32
    uv, vals = O.ref_inputs(20000, N)
    gcf = O.init_gcf(S)
    ref = O.grid_zz(uv, vals, N, gcf, S)
    g = migrid.unpad(migrid.gridGPU(uv, vals, gcf, N, S), N, S)
    assert O.rel_errors(g, ref)[0] < 1e-12


@pytest.mark.parametrize("pair,tol", [("zc", 2e-6), ("cc", 2e-5)])
def test_scatter_float_pairs(pair, tol):
    N, S = 128, 16
    uv, vals = O.gen_uniform(4000, 0, N, seed=7)
    gcf = O.config_gcf(S)
    ref = O.grid_zc(uv, vals, N, gcf, S) if pair == "zc" else O.grid_cc(uv, vals, N, gcf, S)
    g = migrid.unpad(migrid.gridGPU(uv, vals, gcf, N, S, pair=pair), N, S)
    assert O.rel_errors(g, ref)[0] < tol


def test_scatter_planar_layout_and_gridcpu_entry_points():
    N, S = 96, 16
    uv, vals = O.gen_uniform(1500, 0, N, seed=3)
    gcf = O.config_gcf(S)
    ref = O.grid_zz(uv, vals, N, gcf, S)
    out = migrid.gridGPU(uv, vals, gcf, N, S, layout=migrid.LAYOUT_PLANAR)
    assert O.rel_errors(migrid.unpad(out, N, S, planar=True), ref)[0] < 1e-12
    # gridCPU: unpadded planar, zeroed first
    g = migrid.gridCPU(uv, vals, gcf, N, S)
    assert O.rel_errors(g, ref)[0] < 1e-12
    # gridCPU_pz: adds into the planes it is given
    acc = g.copy()
    migrid.gridCPU(uv, vals, gcf, N, S, out=acc)
    assert O.rel_errors(acc, 2 * ref)[0] < 1e-12
    accf = np.zeros((2, N, N), np.float32)
    migrid.gridCPU(uv, vals, gcf, N, S, pair="cc", out=accf)
    assert O.rel_errors(accf, O.grid_cc(uv, vals, N, gcf, S))[0] < 2e-5


def test_scatter_coverage_bit_exact():
    N, S = 128, 16
    uv, vals = O.gen_uniform(3000, -8, N + 8, seed=5)
    ref = O.grid_zz(uv, vals, N, O.config_gcf(S), S, debug1=True)
    out = migrid.gridGPU(uv, vals, O.config_gcf(S), N, S, debug_coverage=True)
    g = migrid.unpad(out, N, S)
    np.testing.assert_array_equal(g.real, ref[0])
    np.testing.assert_array_equal(g.imag, ref[1])


def test_scatter_compute_gcf_compat():
    N, S = 128, 16
    uv, vals = O.gen_uniform(1000, 0, N, seed=9)
    gcf = O.init_gcf(S)
    e = S * (S - 1) // 2 - 1
    r2, i2 = O.compute_gcf_phase(uv, gcf[e, 0], gcf[e, 1])
    ref = O.grid_const_tap(uv, vals, N, S, r2, i2)
    g = migrid.unpad(migrid.gridGPU(uv, vals, gcf, N, S, gcf_mode=migrid.GCF_COMPAT), N, S)
    # sinf/cosf on device vs glibc: a few float ulps in the tap value
    assert O.rel_errors(g, ref)[0] < 1e-6


def test_scatter_analytic_gcf_matches_table():
    N, S = 128, 32
    uv, vals = O.gen_uniform(2000, 0, N, seed=10)
    ref = O.grid_zz(uv, vals, N, O.config_gcf(S, beta=0.05), S)
    g = migrid.unpad(migrid.gridGPU(uv, vals, None, N, S, gcf_mode=migrid.GCF_ANALYTIC, beta=0.05), N, S)
    assert O.rel_errors(g, ref)[0] < 1e-12


def test_scatter_small_and_empty_inputs():
    N, S = 64, 8
    gcf = O.config_gcf(S)
    for npts in (0, 1, 5, 31, 33):
        uv, vals = O.gen_uniform(npts, 0, N, seed=npts)
        ref = O.grid_zz(uv, vals, N, gcf, S)
        g = migrid.unpad(migrid.gridGPU(uv, vals, gcf, N, S), N, S)
        if npts == 0:
            assert not np.abs(g).any()
        else:
            assert O.rel_errors(g, ref)[0] < 1e-12


def test_device_context_accumulates():
    torch = _torch()
    N, S = 128, 16
    uv, vals = O.gen_uniform(3000, 0, N, seed=12)
    gcf = O.config_gcf(S)
    ref = O.grid_zz(uv, vals, N, gcf, S)
    ctx = migrid.Context()
    grid = torch.empty((N, N, 2), dtype=torch.float64, device="cuda")
    ctx.grid(grid, dev(uv), dev(vals), dev(gcf), N, S)
    ctx.grid(grid, dev(uv), dev(vals), dev(gcf), N, S, accumulate=True)
    torch.cuda.synchronize()
    g = grid.cpu().numpy()
    assert O.rel_errors(g[..., 0] + 1j * g[..., 1], 2 * (ref[0] + 1j * ref[1]))[0] < 1e-12


@pytest.mark.parametrize("N,S,H,npts", [(128, 16, 16, 5000), (256, 64, 32, 20000), (200, 32, 32, 7000),
                                        (64, 8, 8, 3000), (1024, 64, 64, 300000)])
def test_bin_sort_bit_exact(N, S, H, npts):
    uv, _ = O.gen_uniform(npts, -S, N + S, seed=N + H)
    perm_ref, offs_ref = O.bin_sort(uv, N, S, H)
    ctx = migrid.Context()
    for dt in ("f64", "f32"):
        t = dev(uv) if dt == "f64" else dev(uv.astype(np.float32))
        perm, offs = ctx.bin_sort(t, N, S, H)
        np.testing.assert_array_equal(perm.cpu().numpy(), perm_ref)
        np.testing.assert_array_equal(offs.cpu().numpy(), offs_ref)


def test_bin_sort_clustered_and_empty():
    N, S, H = 8192, 64, 64
    uv, _ = O.gen_clustered(200000, N, S, mean_r=800.0, seed=77)
    perm_ref, offs_ref = O.bin_sort(uv, N, S, H)
    ctx = migrid.Context()
    perm, offs = ctx.bin_sort(dev(uv), N, S, H)
    np.testing.assert_array_equal(perm.cpu().numpy(), perm_ref)
    np.testing.assert_array_equal(offs.cpu().numpy(), offs_ref)
    perm0, offs0 = ctx.bin_sort(dev(np.zeros((0, 2))), 64, 8, 8)
    assert perm0.numel() == 0 and not offs0.cpu().numpy().any()


def test_bin_sort_dense_keys():
    """40000 visibilities on one key and a few hundred elsewhere: the chunk
    holding that key has more elements than the per-chunk scan takes, so its
    offsets come from the per-key binary search."""
    N, S, H = 256, 16, 16
    rng = np.random.default_rng(5)
    dense = np.tile(np.array([[100.25, 37.5]]), (40000, 1))
    spread = rng.uniform(0, N, (700, 2))
    uv = np.concatenate([spread[:300], dense, spread[300:]])
    perm_ref, offs_ref = O.bin_sort(uv, N, S, H)
    ctx = migrid.Context()
    perm, offs = ctx.bin_sort(dev(uv), N, S, H)
    np.testing.assert_array_equal(perm.cpu().numpy(), perm_ref)
    np.testing.assert_array_equal(offs.cpu().numpy(), offs_ref)
