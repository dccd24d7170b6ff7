"""ctypes wrapper around oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker (or the timed CPU baseline).  The
product library never calls it.

Every function here restates part of the reference; the C source
(oracle/gridcpu.c) cites the reference file:line for each one.  Grids come
back PLANAR, shape (2, N, N) = (real plane, imaginary plane), the layout the
reference gridCPU writes (grid.cu:101-102,153-154).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)


def build() -> str:
    """Compile liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_grid_zz.argtypes = [_dp, _dp, _dp, _i64, _i64, _dp, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_grid_zz_rows.argtypes = [_dp, _dp, _dp, _i64, _i64, _dp, _i64, _i64, _i64]
        L.oracle_grid_zz_omp.argtypes = [_dp, _dp, _dp, _i64, _i64, _dp, _i64, ctypes.c_int]
        L.oracle_grid_zz_omp.restype = ctypes.c_int
        L.oracle_grid_zc.argtypes = [_dp, _fp, _fp, _i64, _i64, _fp, _i64, ctypes.c_int]
        L.oracle_grid_cc.argtypes = [_fp, _fp, _fp, _i64, _i64, _fp, _i64, ctypes.c_int]
        L.oracle_floorf_divergent.argtypes = [_dp, _i64]
        L.oracle_floorf_divergent.restype = _i64
        L.oracle_init_gcf.argtypes = [_dp, _i64]
        L.oracle_config_gcf.argtypes = [_dp, _i64, ctypes.c_double, ctypes.c_double]
        L.oracle_gen_uniform.argtypes = [_dp, _dp, _i64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]
        L.oracle_gen_clustered.argtypes = [_dp, _dp, _i64, _i64, _i64, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_uint64]
        L.oracle_ref_inputs.argtypes = [_dp, _dp, _i64, _i64, ctypes.c_uint]
        L.oracle_vis2ints.argtypes = [_dp, _i64, _ip]
        L.oracle_set_bookmarks.argtypes = [_ip, _i64, _i64, _i64, _ip]
        L.oracle_bin_key.argtypes = [ctypes.c_double, ctypes.c_double, _i64, _i64, _i64]
        L.oracle_bin_key.restype = _i64
        L.oracle_bin_sort.argtypes = [_dp, _i64, _i64, _i64, _i64, _ip, _ip]
        L.oracle_bin_sort.restype = _i64
        L.oracle_window_pixel.argtypes = [_i64, _i64, _i64, _i64]
        L.oracle_window_pixel.restype = _i64
        L.oracle_compute_gcf_phase.argtypes = [_dp, _i64, ctypes.c_double, ctypes.c_double, ctypes.c_int, _fp, _fp]
        L.oracle_grid_const_tap.argtypes = [_dp, _dp, _dp, _i64, _i64, _i64, _fp, _fp]
        L.oracle_grid_const_tap_omp.argtypes = [_dp, _dp, _dp, _i64, _i64, _i64, _fp, _fp,
                                                ctypes.c_int]
        L.oracle_grid_const_tap_omp.restype = ctypes.c_int
        L.oracle_degrid_zz.argtypes = [_dp, _dp, _i64, _dp, _i64, _dp, _i64]
        L.oracle_uvw_to_pixels.argtypes = [_dp, _i64, _dp, _i64, _i64, ctypes.c_double, _i64,
                                           ctypes.c_double, _dp, _ip]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---- gridders -------------------------------------------------------------

def grid_zz(uv, vals, N, gcf, S, floorf_emul=False, out=None, debug1=False):
    """gridCPU (out=None: fresh zeroed grid) / gridCPU_pz (out given:
    accumulate).  uv (n,2) f64, vals (n,2) f64, gcf (64*S*S,2) f64.
    Returns the planar grid (2,N,N) f64."""
    uv = _c(uv, np.float64).reshape(-1, 2)
    vals = _c(vals, np.float64).reshape(-1, 2)
    gcf = _c(gcf, np.float64).reshape(-1, 2)
    zero = out is None
    if out is None:
        out = np.empty((2, N, N), np.float64)
    lib().oracle_grid_zz(_p(out, _dp), _p(uv, _dp), _p(vals, _dp), uv.shape[0], N,
                         _p(gcf, _dp), S, int(floorf_emul), int(zero), int(debug1))
    return out


def grid_zz_omp(uv, vals, N, gcf, S, nthreads):
    """Row-band OpenMP form; bit-identical to grid_zz.  Returns (grid, threads)."""
    uv = _c(uv, np.float64).reshape(-1, 2)
    vals = _c(vals, np.float64).reshape(-1, 2)
    gcf = _c(gcf, np.float64).reshape(-1, 2)
    out = np.empty((2, N, N), np.float64)
    used = lib().oracle_grid_zz_omp(_p(out, _dp), _p(uv, _dp), _p(vals, _dp), uv.shape[0], N,
                                    _p(gcf, _dp), S, int(nthreads))
    return out, used


def grid_zz_rows(uv, vals, N, gcf, S, y_lo, y_hi, out=None):
    uv = _c(uv, np.float64).reshape(-1, 2)
    vals = _c(vals, np.float64).reshape(-1, 2)
    gcf = _c(gcf, np.float64).reshape(-1, 2)
    if out is None:
        out = np.zeros((2, N, N), np.float64)
    lib().oracle_grid_zz_rows(_p(out, _dp), _p(uv, _dp), _p(vals, _dp), uv.shape[0], N,
                              _p(gcf, _dp), S, y_lo, y_hi)
    return out


def grid_zc(uv, vals, N, gcf, S):
    """PRECISION=float, OUTPRECISION=double."""
    uv = _c(uv, np.float32).reshape(-1, 2)
    vals = _c(vals, np.float32).reshape(-1, 2)
    gcf = _c(gcf, np.float32).reshape(-1, 2)
    out = np.empty((2, N, N), np.float64)
    lib().oracle_grid_zc(_p(out, _dp), _p(uv, _fp), _p(vals, _fp), uv.shape[0], N,
                         _p(gcf, _fp), S, 1)
    return out


def grid_cc(uv, vals, N, gcf, S):
    """PRECISION=float, OUTPRECISION=float."""
    uv = _c(uv, np.float32).reshape(-1, 2)
    vals = _c(vals, np.float32).reshape(-1, 2)
    gcf = _c(gcf, np.float32).reshape(-1, 2)
    out = np.empty((2, N, N), np.float32)
    lib().oracle_grid_cc(_p(out, _fp), _p(uv, _fp), _p(vals, _fp), uv.shape[0], N,
                         _p(gcf, _fp), S, 1)
    return out


def floorf_divergent(uv):
    uv = _c(uv, np.float64).reshape(-1, 2)
    return int(lib().oracle_floorf_divergent(_p(uv, _dp), uv.shape[0]))


# ---- GCF tables -------------------------------------------------------------

def init_gcf(S):
    """The reference's init_gcf table, (64*S*S, 2) f64."""
    g = np.empty((64 * S * S, 2), np.float64)
    lib().oracle_init_gcf(_p(g, _dp), S)
    return g


def config_gcf(S, sigma=None, beta=0.05):
    """Gaussian (sigma = S/6 by default) x exp(i*beta*r^2), (64*S*S, 2) f64."""
    if sigma is None:
        sigma = S / 6.0
    g = np.empty((64 * S * S, 2), np.float64)
    lib().oracle_config_gcf(_p(g, _dp), S, float(sigma), float(beta))
    return g


# ---- generators -------------------------------------------------------------

def gen_uniform(npts, lo, hi, seed):
    uv = np.empty((npts, 2), np.float64)
    vals = np.empty((npts, 2), np.float64)
    lib().oracle_gen_uniform(_p(uv, _dp), _p(vals, _dp), npts, float(lo), float(hi), int(seed))
    return uv, vals


def gen_clustered(npts, N, S, mean_r=800.0, seed=1541618, cx=None, cy=None):
    cx = N / 2.0 if cx is None else cx
    cy = N / 2.0 if cy is None else cy
    uv = np.empty((npts, 2), np.float64)
    vals = np.empty((npts, 2), np.float64)
    lib().oracle_gen_clustered(_p(uv, _dp), _p(vals, _dp), npts, N, S, float(mean_r),
                               float(cx), float(cy), int(seed))
    return uv, vals


def ref_inputs(npts, img_size, seed=1541617):
    uv = np.empty((npts, 2), np.float64)
    vals = np.empty((npts, 2), np.float64)
    lib().oracle_ref_inputs(_p(uv, _dp), _p(vals, _dp), npts, img_size, seed)
    return uv, vals


# ---- binning ----------------------------------------------------------------

def vis2ints(uv):
    uv = _c(uv, np.float64).reshape(-1, 2)
    out = np.empty((uv.shape[0], 2), np.int32)
    lib().oracle_vis2ints(_p(uv, _dp), uv.shape[0], _p(out, _ip))
    return out


def set_bookmarks(ints, blocksize, blockgrid):
    ints = _c(ints, np.int32).reshape(-1, 2)
    bm = np.empty(blockgrid * blockgrid + 1, np.int32)
    lib().oracle_set_bookmarks(_p(ints, _ip), ints.shape[0], blocksize, blockgrid, _p(bm, _ip))
    return bm


def n_bins(N, S, H):
    return (N + S - 2) // H + 1


def bin_key(u, v, N, S, H):
    return int(lib().oracle_bin_key(float(u), float(v), N, S, H))


def bin_sort(uv, N, S, H):
    """Returns (perm, offsets): stable order of visibility indices by
    (k, bin_y, bin_x) key and the CSR offsets (length 64*nb*nb + 1)."""
    uv = _c(uv, np.float64).reshape(-1, 2)
    nb = n_bins(N, S, H)
    perm = np.empty(max(uv.shape[0], 1), np.int32)
    offs = np.empty(64 * nb * nb + 1, np.int32)
    total = lib().oracle_bin_sort(_p(uv, _dp), uv.shape[0], N, S, H, _p(perm, _ip), _p(offs, _ip))
    return perm[:total].copy(), offs


def window_pixel(m, t, S, N):
    return int(lib().oracle_window_pixel(m, t, S, N))


def compute_gcf_phase(uv, T, w, fixed_point=False):
    """COMPUTE_GCF tap values (r2, i2) per visibility; fixed_point selects the
    gather/moving-window form of dx."""
    uv = _c(uv, np.float64).reshape(-1, 2)
    r2 = np.empty(uv.shape[0], np.float32)
    i2 = np.empty(uv.shape[0], np.float32)
    lib().oracle_compute_gcf_phase(_p(uv, _dp), uv.shape[0], float(T), float(w), int(fixed_point),
                                   _p(r2, _fp), _p(i2, _fp))
    return r2, i2


def grid_const_tap(uv, vals, N, S, r2, i2):
    uv = _c(uv, np.float64).reshape(-1, 2)
    vals = _c(vals, np.float64).reshape(-1, 2)
    r2 = _c(r2, np.float32)
    i2 = _c(i2, np.float32)
    out = np.empty((2, N, N), np.float64)
    lib().oracle_grid_const_tap(_p(out, _dp), _p(uv, _dp), _p(vals, _dp), uv.shape[0], N, S,
                                _p(r2, _fp), _p(i2, _fp))
    return out


def grid_const_tap_omp(uv, vals, N, S, r2, i2, nthreads):
    """Row-band OpenMP form of grid_const_tap; bit-identical to it.
    Returns (grid, threads)."""
    uv = _c(uv, np.float64).reshape(-1, 2)
    vals = _c(vals, np.float64).reshape(-1, 2)
    r2 = _c(r2, np.float32)
    i2 = _c(i2, np.float32)
    out = np.empty((2, N, N), np.float64)
    used = lib().oracle_grid_const_tap_omp(_p(out, _dp), _p(uv, _dp), _p(vals, _dp), uv.shape[0],
                                           N, S, _p(r2, _fp), _p(i2, _fp), int(nthreads))
    return out, used


def degrid_zz(uv, grid, N, gcf, S):
    """Adjoint of grid_zz: grid planar (2,N,N) -> values (n,2)."""
    uv = _c(uv, np.float64).reshape(-1, 2)
    grid = _c(grid, np.float64).reshape(2, N, N)
    gcf = _c(gcf, np.float64).reshape(-1, 2)
    vals = np.empty((uv.shape[0], 2), np.float64)
    lib().oracle_degrid_zz(_p(vals, _dp), _p(uv, _dp), uv.shape[0], _p(grid, _dp), N,
                           _p(gcf, _dp), S)
    return vals


def uvw_to_pixels(uvw, freq, N, cell, nplanes=0, w_max=1.0):
    """(nrow,3) metres, (nchan,) Hz -> (nrow*nchan, 2) pixels and, with
    nplanes > 0, the w-plane index per visibility."""
    uvw = _c(uvw, np.float64).reshape(-1, 3)
    freq = _c(freq, np.float64).reshape(-1)
    n = uvw.shape[0] * freq.shape[0]
    uv = np.empty((n, 2), np.float64)
    g = np.empty(n, np.int32) if nplanes > 0 else None
    lib().oracle_uvw_to_pixels(_p(uvw, _dp), uvw.shape[0], _p(freq, _dp), freq.shape[0], N,
                               float(cell), int(nplanes), float(w_max), _p(uv, _dp),
                               _p(g, _ip) if g is not None else None)
    return uv, g


# ---- grid <-> image (no reference counterpart; numpy's FFT) -----------------

def _taper2d(taper, N):
    if taper is None:
        return 1.0
    t = np.asarray(taper, np.float64).reshape(N)
    return t[:, None] * t[None, :]


def grid_to_image(grid, scale=1.0, taper=None):
    """Centred unnormalised transform of a complex (N, N) grid (rows v, columns
    u) to the image plane:
      img[y][x] = scale / (taper[x] taper[y]) *
                  sum grid[v][u] exp(+2 pi i ((u - N/2)(x - N/2) + (v - N/2)(y - N/2)) / N)
    = fftshift(ifft2(ifftshift(grid))) * N^2 * scale / taper2d (even N)."""
    g = np.asarray(grid, np.complex128)
    N = g.shape[0]
    img = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(g))) * (N * N)
    return img * scale / _taper2d(taper, N)


def image_to_grid(img, scale=1.0, taper=None):
    """Adjoint of grid_to_image: the image divided by the taper, then the
    exp(-...) centred transform, times scale."""
    a = np.asarray(img, np.complex128)
    N = a.shape[0]
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(a / _taper2d(taper, N)))) * scale


def centred_dft_direct(a, sign):
    """The centred 2-D DFT by its definition, O(N^4): tiny N only (pins the
    shift conventions of grid_to_image / image_to_grid)."""
    a = np.asarray(a, np.complex128)
    N = a.shape[0]
    k = np.arange(N) - N // 2
    E = np.exp(sign * 2j * np.pi * np.outer(k, k) / N)  # E[x, u]
    return E @ a @ E.T


# ---- error metrics ----------------------------------------------------------

def rel_errors(got, ref):
    """(rel_max, rel_rms) over complex moduli: max|dG|/max|G_ref| and
    ||dG||_2/||G_ref||_2.  Both arguments planar (2,N,N) or complex (N,N)."""
    def cplx(a):
        a = np.asarray(a)
        if np.iscomplexobj(a):
            return a.astype(np.complex128)
        return a[0].astype(np.float64) + 1j * a[1].astype(np.float64)
    g = cplx(got)
    r = cplx(ref)
    d = np.abs(g - r)
    m = np.abs(r)
    rmax = float(d.max() / m.max()) if m.max() > 0 else float(d.max())
    rrms = float(np.sqrt((d ** 2).sum() / (m ** 2).sum())) if m.max() > 0 else float(np.sqrt((d ** 2).sum()))
    return rmax, rrms
