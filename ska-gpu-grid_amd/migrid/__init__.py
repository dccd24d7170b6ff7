"""Python binding over the libmigrid C ABI (include/migrid.h).

Replaces the reference's Python-2 extension GPUGrid.convgrid
(GPUGrid.cpp:124-164) with a ctypes binding: numpy arrays in, the padded or
unpadded grid out, errors raised instead of printed (GPUGrid.cpp:128-141),
and a correctly sized output (the reference allocates 16*npts bytes for an
img_size^2 grid, GPUGrid.cpp:148,159-160).

Device-resident calls take torch tensors (device memory and streams are
torch's; the computation is the HIP library's).  There is no CPU fallback:
if libmigrid.so is missing or cannot load, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmigrid.so")

SCATTER, GATHER, WINDOW, TILED = 0, 1, 2, 3
LAYOUT_DEFAULT, LAYOUT_INTERLEAVED, LAYOUT_PLANAR = 0, 1, 2
GCF_TABLE, GCF_COMPAT, GCF_ANALYTIC = 0, 1, 2
ARITH_NATIVE, ARITH_F32, ARITH_F32_BF16GCF = 0, 1, 2
F64, F32 = 0, 1
FFT_TO_IMAGE, FFT_TO_GRID = 1, -1

STRATEGIES = {"scatter": SCATTER, "gather": GATHER, "window": WINDOW, "tiled": TILED}

_ERRNAMES = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ENOTSUP", -5: "EBUSY"}


class MigridError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"migrid {_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [
        ("strategy", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("gcf_mode", ctypes.c_int32),
        ("arith", ctypes.c_int32),
        ("accumulate", ctypes.c_int32),
        ("debug_coverage", ctypes.c_int32),
        ("tile", ctypes.c_int32),
        ("vis_per_wave", ctypes.c_int32),
        ("gcf_sigma", ctypes.c_double),
        ("gcf_beta", ctypes.c_double),
        ("stage_min", ctypes.c_int32),
        ("wg_per_cu", ctypes.c_int32),
        ("tap_smax", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 5),
    ]


class Job(ctypes.Structure):
    _fields_ = [
        ("in_type", ctypes.c_int32),
        ("out_type", ctypes.c_int32),
        ("uv", ctypes.c_void_p),
        ("vals", ctypes.c_void_p),
        ("gcf", ctypes.c_void_p),
        ("grid", ctypes.c_void_p),
        ("grid_lo", ctypes.c_void_p),
        ("npts", ctypes.c_int64),
        ("img_dim", ctypes.c_int64),
        ("gcf_dim", ctypes.c_int64),
    ]


class ServerCfg(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("in_type", ctypes.c_int32),
        ("out_type", ctypes.c_int32),
        ("max_batch", ctypes.c_int32),
        ("max_inflight", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("hbm_budget", ctypes.c_int64),
        ("opts", Opts),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "submitted", "completed", "failed", "batches", "vis_completed", "hbm_in_use", "hbm_budget",
        "jobs_resident", "queued", "running")] + [
        ("gpu_ms", ctypes.c_double), ("gpu_busy_ms", ctypes.c_double),
        ("overlap_ms", ctypes.c_double), ("sorts_reused", ctypes.c_int64)]


class Ceiling(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("ms", "amount", "per_s", "per_clk_cu", "clock_ghz")]


def ceiling_lds(elem_bytes, pitch, iters=20000):
    """LDS read bandwidth probe (migrid_ceiling_lds) as a dict; pitch 0 reads
    linear rows, else the owner kernel's lane block at that row pitch."""
    c = Ceiling()
    check(lib().migrid_ceiling_lds(elem_bytes, pitch, iters, ctypes.byref(c)))
    return {n: getattr(c, n) for n, _ in Ceiling._fields_}


def ceiling_atomics(grid, uv, N, S):
    """The scatter kernel's atomics alone (migrid_ceiling_atomics) as a dict."""
    j = make_job(grid, uv, uv, None, N, S)
    c = Ceiling()
    check(lib().migrid_ceiling_atomics(ctypes.byref(j), ctypes.byref(c)))
    return {n: getattr(c, n) for n, _ in Ceiling._fields_}


# Every symbol include/migrid.h declares (checked by the CPU test suite).
EXPORTS = [
    "migrid_opts_default", "migrid_last_error", "migrid_version",
    "migrid_gridGPU_zz", "migrid_gridGPU_zc", "migrid_gridGPU_cc",
    "migrid_gridCPU_zz", "migrid_gridCPU_pz_zz", "migrid_gridCPU_pz_zc", "migrid_gridCPU_pz_cc",
    "migrid_padded_elems",
    "migrid_ctx_create", "migrid_ctx_destroy", "migrid_ctx_grid",
    "migrid_ctx_bin_sort", "migrid_bins_per_axis",
    "migrid_ctx_set_profiling", "migrid_ctx_profile_read", "migrid_ctx_profile_reset",
    "migrid_batch_create", "migrid_batch_run", "migrid_batch_graph_launch", "migrid_batch_info",
    "migrid_batch_destroy", "migrid_autotune",
    "migrid_server_cfg_default", "migrid_server_create", "migrid_server_submit", "migrid_server_poll",
    "migrid_server_wait", "migrid_server_read_grid", "migrid_server_drop_job",
    "migrid_server_get_stats", "migrid_server_destroy",
    "migrid_uvw_to_pixels", "migrid_ctx_grid_indexed", "migrid_read_uvw_text",
    "migrid_ctx_degrid", "migrid_ctx_fft", "migrid_grid_noalloc",
    "migrid_ceiling_lds", "migrid_ceiling_atomics",
]

_lib = None


def lib():
    """Load libmigrid.so (raises if it is missing: no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MigridError(-3, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i64, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int64, ctypes.c_int32
    L.migrid_opts_default.argtypes = [ctypes.POINTER(Opts)]
    L.migrid_opts_default.restype = None
    L.migrid_last_error.restype = ctypes.c_char_p
    L.migrid_version.restype = ctypes.c_char_p
    for name in ("migrid_gridGPU_zz", "migrid_gridGPU_zc", "migrid_gridGPU_cc"):
        f = getattr(L, name)
        f.argtypes = [vp, vp, vp, sz, sz, vp, sz, ctypes.POINTER(Opts)]
        f.restype = ctypes.c_int
    for name in ("migrid_gridCPU_zz", "migrid_gridCPU_pz_zz", "migrid_gridCPU_pz_zc",
                 "migrid_gridCPU_pz_cc"):
        f = getattr(L, name)
        f.argtypes = [vp, vp, vp, sz, sz, vp, sz]
        f.restype = ctypes.c_int
    L.migrid_padded_elems.argtypes = [sz, sz]
    L.migrid_padded_elems.restype = sz
    L.migrid_ctx_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    L.migrid_ctx_destroy.argtypes = [vp]
    L.migrid_ctx_grid.argtypes = [vp, ctypes.POINTER(Job), ctypes.POINTER(Opts), vp]
    L.migrid_ctx_bin_sort.argtypes = [vp, i32, vp, i64, i64, i64, i32, vp, vp,
                                      ctypes.POINTER(i64), vp]
    L.migrid_ctx_set_profiling.argtypes = [vp, ctypes.c_int]
    L.migrid_ctx_profile_read.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.migrid_ctx_profile_reset.argtypes = [vp]
    L.migrid_bins_per_axis.argtypes = [i64, i64, i32]
    L.migrid_bins_per_axis.restype = i64
    L.migrid_batch_create.argtypes = [vp, ctypes.POINTER(Job), ctypes.c_int, ctypes.POINTER(Opts),
                                      ctypes.POINTER(vp)]
    L.migrid_batch_run.argtypes = [vp, vp]
    L.migrid_batch_graph_launch.argtypes = [vp, vp]
    L.migrid_batch_info.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.migrid_batch_destroy.argtypes = [vp]
    L.migrid_autotune.argtypes = [vp, ctypes.POINTER(Job), ctypes.c_int, ctypes.POINTER(Opts),
                                  ctypes.POINTER(Opts), ctypes.POINTER(ctypes.c_double), vp]
    L.migrid_server_cfg_default.argtypes = [ctypes.POINTER(ServerCfg)]
    L.migrid_server_cfg_default.restype = None
    L.migrid_server_create.argtypes = [ctypes.POINTER(ServerCfg), ctypes.POINTER(vp)]
    L.migrid_server_submit.argtypes = [vp, i64, i64, i64, vp, vp, vp, i64, i32, ctypes.POINTER(i64)]
    L.migrid_server_poll.argtypes = [vp, i64, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double)]
    L.migrid_server_wait.argtypes = [vp, i64, ctypes.POINTER(ctypes.c_double)]
    L.migrid_server_read_grid.argtypes = [vp, i64, vp]
    L.migrid_server_drop_job.argtypes = [vp, i64]
    L.migrid_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.migrid_server_destroy.argtypes = [vp]
    L.migrid_uvw_to_pixels.argtypes = [vp, i64, vp, i32, i64, ctypes.c_double, i32, ctypes.c_double,
                                       vp, vp, vp]
    L.migrid_ctx_grid_indexed.argtypes = [vp, ctypes.POINTER(Job), i32, vp, ctypes.POINTER(Opts), vp]
    L.migrid_read_uvw_text.argtypes = [ctypes.c_char_p, i64, i64, vp, vp, ctypes.POINTER(i64)]
    L.migrid_ctx_degrid.argtypes = [vp, ctypes.POINTER(Job), vp, ctypes.POINTER(Opts), vp]
    L.migrid_ctx_fft.argtypes = [vp, i32, i32, vp, vp, i64, i32, ctypes.c_double, vp, vp]
    L.migrid_grid_noalloc.argtypes = [ctypes.POINTER(Job), ctypes.POINTER(Opts), vp]
    L.migrid_ceiling_lds.argtypes = [i32, i32, i32, ctypes.POINTER(Ceiling)]
    L.migrid_ceiling_atomics.argtypes = [ctypes.POINTER(Job), ctypes.POINTER(Ceiling)]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise MigridError(rc, lib().migrid_last_error().decode())


def version():
    return lib().migrid_version().decode()


def make_opts(strategy="scatter", layout=LAYOUT_DEFAULT, gcf_mode=GCF_TABLE, arith=ARITH_NATIVE,
              accumulate=False, debug_coverage=False, tile=0, vis_per_wave=0, sigma=0.0,
              beta=0.05, stage_min=0, wg_per_cu=0, tap_smax=0):
    """migrid_opts from keywords (fields as in include/migrid.h).  Tiled
    tuning: tile edge, visibilities per work item, staging threshold,
    workgroups per CU, and tap_smax, the largest support the tap kernel grids
    (0: 16, -1: never)."""
    o = Opts()
    lib().migrid_opts_default(ctypes.byref(o))
    o.stage_min = stage_min
    o.wg_per_cu = wg_per_cu
    o.tap_smax = tap_smax
    o.strategy = STRATEGIES[strategy] if isinstance(strategy, str) else int(strategy)
    o.layout = layout
    o.gcf_mode = gcf_mode
    o.arith = arith
    o.accumulate = int(bool(accumulate))
    o.debug_coverage = int(bool(debug_coverage))
    o.tile = tile
    o.vis_per_wave = vis_per_wave
    o.gcf_sigma = sigma
    o.gcf_beta = beta
    return o


def padded_elems(N, S):
    return int(lib().migrid_padded_elems(N, S))


def grid_base(N, S):
    """Offset (complex elements) of the grid base inside the padded buffer."""
    return N * S + S


# ---- host (drop-in) calls -----------------------------------------------------

def gridGPU(uv, vals, gcf, N, S, pair="zz", out=None, **kw):
    """gridGPU<TO,T> (grid_gpu.cu:675-819) on numpy arrays.

    uv, vals: (n,2); gcf: (64*S*S,2); pair in {"zz","zc","cc"} picks the
    reference instantiation <double2,double2>, <double2,float2>,
    <float2,float2>.  Returns the padded buffer as a (P,2) array (P =
    padded_elems(N,S)); use unpad() for the grid."""
    ti = np.float64 if pair == "zz" else np.float32
    to = np.float32 if pair == "cc" else np.float64
    uv = np.ascontiguousarray(uv, ti).reshape(-1, 2)
    vals = np.ascontiguousarray(vals, ti).reshape(-1, 2)
    gcf_p = None
    if gcf is not None:
        gcf = np.ascontiguousarray(gcf, ti).reshape(-1, 2)
        if gcf.shape[0] != 64 * S * S:
            raise MigridError(-1, f"gcf has {gcf.shape[0]} elements, expected 64*S*S={64 * S * S}")
        gcf_p = gcf.ctypes.data
    if uv.shape != vals.shape:
        raise MigridError(-1, "uv and vals differ in length")
    P = padded_elems(N, S)
    if out is None:
        out = np.zeros((P, 2), to)
    elif out.shape != (P, 2) or out.dtype != to or not out.flags.c_contiguous:
        raise MigridError(-1, f"out must be a contiguous ({P},2) {np.dtype(to).name} array")
    o = make_opts(**kw)
    f = getattr(lib(), f"migrid_gridGPU_{pair}")
    check(f(out.ctypes.data, uv.ctypes.data, vals.ctypes.data, uv.shape[0], N, gcf_p, S,
            ctypes.byref(o)))
    return out


def gridCPU(uv, vals, gcf, N, S, pair="zz", out=None):
    """gridCPU / gridCPU_pz (grid.cu:87-193) computed on the GPU.  Returns
    the planar grid (2,N,N).  With out=None the grid is fresh (gridCPU);
    with out given the call adds into it (gridCPU_pz)."""
    ti = np.float64 if pair == "zz" else np.float32
    to = np.float32 if pair == "cc" else np.float64
    uv = np.ascontiguousarray(uv, ti).reshape(-1, 2)
    vals = np.ascontiguousarray(vals, ti).reshape(-1, 2)
    gcf = np.ascontiguousarray(gcf, ti).reshape(-1, 2)
    if out is None:
        out = np.zeros((2, N, N), to)
        name = "migrid_gridCPU_zz" if pair == "zz" else f"migrid_gridCPU_pz_{pair}"
    else:
        if out.shape != (2, N, N) or out.dtype != to or not out.flags.c_contiguous:
            raise MigridError(-1, "out must be a contiguous (2,N,N) array of the output type")
        name = f"migrid_gridCPU_pz_{pair}"
    check(getattr(lib(), name)(out.ctypes.data, uv.ctypes.data, vals.ctypes.data, uv.shape[0], N,
                               gcf.ctypes.data, S))
    return out


def unpad(padded, N, S, planar=False):
    """Grid view of a padded buffer: complex (N,N) for interleaved, or the
    planar (2,N,N) real/imaginary planes."""
    base = grid_base(N, S)
    flat = padded.reshape(-1)
    if planar:
        return flat[2 * base: 2 * base + 2 * N * N].reshape(2, N, N)
    g = padded[base: base + N * N]
    return (g[:, 0] + 1j * g[:, 1]).reshape(N, N)


# ---- device-resident calls (torch tensors) ------------------------------------

class Context:
    """A migrid_ctx: device scratch reused across calls."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().migrid_ctx_create(ctypes.byref(h), device))
        self.h = h

    def close(self):
        if self.h:
            lib().migrid_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self, grid, uv, vals, gcf, N, S, grid_lo=None, stream=None, **kw):
        """Grid on device tensors.  grid: unpadded (N*N*2,) or (N,N,2) /
        (2,N,N) tensor of float64/float32; uv, vals: (n,2); gcf: (64*S*S,2)
        or None (ANALYTIC).  Asynchronous on `stream` (default: torch's
        current stream)."""
        import torch
        j = Job()
        j.in_type = F64 if uv.dtype == torch.float64 else F32
        j.out_type = F64 if grid.dtype == torch.float64 else F32
        for t in (uv, vals, grid) + ((gcf,) if gcf is not None else ()):
            if not t.is_contiguous() or not t.is_cuda:
                raise MigridError(-1, "device tensors must be contiguous and on the GPU")
        j.uv = uv.data_ptr()
        j.vals = vals.data_ptr()
        j.gcf = gcf.data_ptr() if gcf is not None else None
        j.grid = grid.data_ptr()
        j.grid_lo = grid_lo.data_ptr() if grid_lo is not None else None
        j.npts = uv.shape[0]
        j.img_dim = N
        j.gcf_dim = S
        o = make_opts(**kw)
        if stream is None:
            stream = torch.cuda.current_stream()
        check(lib().migrid_ctx_grid(self.h, ctypes.byref(j), ctypes.byref(o),
                                    ctypes.c_void_p(stream.cuda_stream)))

    def profiling(self, on=True):
        """Record HIP events around the main gridding kernel of each call."""
        check(lib().migrid_ctx_set_profiling(self.h, int(bool(on))))
        check(lib().migrid_ctx_profile_reset(self.h))

    def kernel_ms(self, cap=1 << 16):
        """Per-launch durations (ms) of the main kernel since profiling()."""
        buf = (ctypes.c_double * cap)()
        n = lib().migrid_ctx_profile_read(self.h, buf, cap)
        if n < 0:
            check(n)
        return [buf[i] for i in range(min(n, cap))]

    def degrid(self, grid, uv, gcf, N, S, out=None, stream=None, **kw):
        """Degridding (the adjoint of grid): visibility values (n, 2) in uv's
        dtype from the (N,N,2) or planar (2,N,N) grid.  Asynchronous."""
        import torch
        if out is None:
            out = torch.empty((uv.shape[0], 2), dtype=uv.dtype, device=uv.device)
        j = make_job(grid, uv, out, gcf, N, S)
        o = make_opts(**kw)
        s = ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)
        check(lib().migrid_ctx_degrid(self.h, ctypes.byref(j), out.data_ptr(), ctypes.byref(o), s))
        return out

    def fft(self, x, out=None, direction=FFT_TO_IMAGE, scale=1.0, taper=None, planar=False,
            stream=None):
        """Centred unnormalised 2-D FFT between grid and image: FFT_TO_IMAGE
        gives fftshift(ifft2(ifftshift(x))) * N^2 * scale / (taper[x] taper[y]),
        FFT_TO_GRID its adjoint.  x: (N,N,2) tensor of float64/float32, or
        (2,N,N) with planar=True; taper: None or (N,) float64 device tensor.
        Returns the (N,N,2) result (out may be x itself when interleaved)."""
        import torch
        N = x.shape[1] if planar else x.shape[0]
        if out is None:
            out = torch.empty((N, N, 2), dtype=x.dtype, device=x.device)
        for t in (x, out) + ((taper,) if taper is not None else ()):
            if not t.is_contiguous() or not t.is_cuda:
                raise MigridError(-1, "device tensors must be contiguous and on the GPU")
        if taper is not None and (taper.dtype != torch.float64 or taper.numel() != N):
            raise MigridError(-1, "taper must be N float64 values")
        typ = F64 if x.dtype == torch.float64 else F32
        s = ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)
        check(lib().migrid_ctx_fft(self.h, typ, LAYOUT_PLANAR if planar else LAYOUT_INTERLEAVED,
                                   x.data_ptr(), out.data_ptr(), N, direction, float(scale),
                                   taper.data_ptr() if taper is not None else None, s))
        return out

    def grid_indexed(self, grid, uv, vals, gcfs, gcf_index, N, S, stream=None, **kw):
        """Gridding with a GCF table per visibility: gcfs (ngcf*64*S*S, 2) or
        (ngcf, 64*S*S, 2), gcf_index (n,) int32 device tensor."""
        import torch
        ngcf = gcfs.numel() // (64 * S * S * 2)
        if ngcf * 64 * S * S * 2 != gcfs.numel():
            raise MigridError(-1, "gcfs must hold a whole number of 64*S*S tables")
        if gcf_index.dtype != torch.int32 or not gcf_index.is_cuda or not gcf_index.is_contiguous():
            raise MigridError(-1, "gcf_index must be a contiguous int32 device tensor")
        j = make_job(grid, uv, vals, gcfs, N, S)
        o = make_opts(**{"strategy": "tiled", **kw})
        s = ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)
        check(lib().migrid_ctx_grid_indexed(self.h, ctypes.byref(j), ngcf, gcf_index.data_ptr(),
                                            ctypes.byref(o), s))

    def bin_sort(self, uv, N, S, bin_size, stream=None):
        """Device bin sort; returns (perm, offsets) torch int32 tensors."""
        import torch
        nb = int(lib().migrid_bins_per_axis(N, S, bin_size))
        n = uv.shape[0]
        perm = torch.empty(max(n, 1), dtype=torch.int32, device=uv.device)
        offs = torch.empty(64 * nb * nb + 1, dtype=torch.int32, device=uv.device)
        cnt = ctypes.c_int64()
        if stream is None:
            stream = torch.cuda.current_stream()
        check(lib().migrid_ctx_bin_sort(self.h, F64 if uv.dtype == torch.float64 else F32,
                                        uv.data_ptr(), n, N, S, bin_size, perm.data_ptr(),
                                        offs.data_ptr(), ctypes.byref(cnt),
                                        ctypes.c_void_p(stream.cuda_stream)))
        return perm[: cnt.value], offs


def read_uvw_text(path, N, max_rows=None):
    """The reference's UVW text format -> (uv (n,2) pixels, w (n,)) numpy."""
    if max_rows is None:
        with open(path) as f:
            max_rows = sum(1 for _ in f)
    uv = np.empty((max(max_rows, 1), 2), np.float64)
    w = np.empty(max(max_rows, 1), np.float64)
    n = ctypes.c_int64()
    check(lib().migrid_read_uvw_text(os.fsencode(path), N, max_rows, uv.ctypes.data, w.ctypes.data,
                                     ctypes.byref(n)))
    return uv[: n.value], w[: n.value]


def uvw_to_pixels(uvw, freq, N, cell, nplanes=0, w_max=1.0, stream=None):
    """Device tensors uvw (nrow,3) metres and freq (nchan,) Hz (float64) ->
    (uv (nrow*nchan, 2) pixels, w-plane index (nrow*nchan,) int32 or None)."""
    import torch
    for t in (uvw, freq):
        if not t.is_cuda or not t.is_contiguous() or t.dtype != torch.float64:
            raise MigridError(-1, "uvw and freq must be contiguous float64 device tensors")
    nrow, nchan = uvw.shape[0], freq.shape[0]
    uv = torch.empty((nrow * nchan, 2), dtype=torch.float64, device=uvw.device)
    g = torch.empty(nrow * nchan, dtype=torch.int32, device=uvw.device) if nplanes > 0 else None
    s = ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)
    check(lib().migrid_uvw_to_pixels(uvw.data_ptr(), nrow, freq.data_ptr(), nchan, N, float(cell),
                                     nplanes, float(w_max), uv.data_ptr(),
                                     g.data_ptr() if g is not None else None, s))
    return uv, g


def grid_noalloc(grid, uv, vals, gcf, N, S, grid_lo=None, stream=None, **kw):
    """migrid_grid_noalloc on device tensors: Context.grid's work with the
    device scratch taken from the library's static per-GPU array (no device
    allocation; jobs beyond its size raise).  Asynchronous on `stream`."""
    import torch
    for t in (uv, vals, grid) + ((gcf,) if gcf is not None else ()):
        if not t.is_contiguous() or not t.is_cuda:
            raise MigridError(-1, "device tensors must be contiguous and on the GPU")
    j = make_job(grid, uv, vals, gcf, N, S, grid_lo)
    o = make_opts(**kw)
    if stream is None:
        stream = torch.cuda.current_stream()
    check(lib().migrid_grid_noalloc(ctypes.byref(j), ctypes.byref(o),
                                    ctypes.c_void_p(stream.cuda_stream)))


def make_job(grid, uv, vals, gcf, N, S, grid_lo=None):
    """A Job structure over device tensors (see Context.grid for shapes)."""
    import torch
    j = Job()
    j.in_type = F64 if uv.dtype == torch.float64 else F32
    j.out_type = F64 if grid.dtype == torch.float64 else F32
    for t in (uv, vals, grid) + ((gcf,) if gcf is not None else ()):
        if not t.is_contiguous() or not t.is_cuda:
            raise MigridError(-1, "device tensors must be contiguous and on the GPU")
    j.uv = uv.data_ptr()
    j.vals = vals.data_ptr()
    j.gcf = gcf.data_ptr() if gcf is not None else None
    j.grid = grid.data_ptr()
    j.grid_lo = grid_lo.data_ptr() if grid_lo is not None else None
    j.npts = uv.shape[0]
    j.img_dim = N
    j.gcf_dim = S
    return j


class Batch:
    """Many jobs gridded by one persistent kernel launch (optionally replayed
    from a captured hipGraph).  jobs: list of dicts with keys grid, uv, vals,
    gcf, N, S (and optional grid_lo); the tensors must outlive the batch."""

    def __init__(self, ctx, jobs, **opts):
        self._keep = jobs
        arr = (Job * len(jobs))(*[make_job(j["grid"], j["uv"], j["vals"], j.get("gcf"), j["N"], j["S"],
                                           j.get("grid_lo")) for j in jobs])
        self._jobs = arr
        o = opts.pop("opts", None) or make_opts(**{"strategy": "tiled", **opts})
        h = ctypes.c_void_p()
        check(lib().migrid_batch_create(ctx.h, arr, len(jobs), ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        self.ctx = ctx

    def _stream(self, stream):
        import torch
        return ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)

    def run(self, stream=None):
        check(lib().migrid_batch_run(self.h, self._stream(stream)))

    def graph_launch(self, stream=None):
        check(lib().migrid_batch_graph_launch(self.h, self._stream(stream)))

    def info(self):
        a, t, n = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().migrid_batch_info(self.h, ctypes.byref(a), ctypes.byref(t), ctypes.byref(n)))
        return {"arena_bytes": a.value, "tile": t.value, "njobs": n.value}

    def close(self):
        if getattr(self, "h", None):
            lib().migrid_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(ctx, jobs, stream=None, **opts):
    """Returns (Opts, ms) of the fastest tile / work-item configuration."""
    import torch
    arr = (Job * len(jobs))(*[make_job(j["grid"], j["uv"], j["vals"], j.get("gcf"), j["N"], j["S"],
                                       j.get("grid_lo")) for j in jobs])
    base = make_opts(**{"strategy": "tiled", **opts})
    best = Opts()
    ms = ctypes.c_double()
    s = ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)
    check(lib().migrid_autotune(ctx.h, arr, len(jobs), ctypes.byref(base), ctypes.byref(best),
                                ctypes.byref(ms), s))
    return best, ms.value


# ---- serving loop -------------------------------------------------------------

SUBMIT_HOST, SUBMIT_DEVICE, SUBMIT_KEEP_SORT, SUBMIT_SAME_UV = 0, 1, 2, 4


class Server:
    """Continuous-batching request loop over resident per-job grids.

    submit() takes numpy arrays (copied at submit) or device torch tensors
    (which must stay alive until the request completes); read_grid() returns
    the job's accumulated N x N complex grid as numpy."""

    def __init__(self, in_type=F64, out_type=F64, max_batch=0, max_inflight=0, hbm_budget=0,
                 device=0, **opts):
        c = ServerCfg()
        lib().migrid_server_cfg_default(ctypes.byref(c))
        c.device = device
        c.in_type, c.out_type = in_type, out_type
        c.max_batch, c.max_inflight, c.hbm_budget = max_batch, max_inflight, int(hbm_budget)
        c.opts = make_opts(**{"strategy": "tiled", "arith": ARITH_F32, **opts})
        self.cfg = c
        h = ctypes.c_void_p()
        check(lib().migrid_server_create(ctypes.byref(c), ctypes.byref(h)))
        self.h = h
        self._keep = {}
        self._N = {}

    def submit(self, job_id, uv, vals, gcf, N, S, keep_sort=False, same_uv=False):
        """Queues one request.  keep_sort: keep its bin sort resident with the
        job; same_uv: its uv are those of the job's kept sort (the sort is
        reused, only the values are packed)."""
        extra = (SUBMIT_KEEP_SORT if keep_sort else 0) | (SUBMIT_SAME_UV if same_uv else 0)
        ctype = np.complex128 if self.cfg.in_type == F64 else np.complex64
        rtype = np.float64 if self.cfg.in_type == F64 else np.float32
        t = ctypes.c_int64()
        if isinstance(uv, np.ndarray):
            uv = np.ascontiguousarray(uv, dtype=rtype).reshape(-1, 2)
            vals = np.ascontiguousarray(vals)
            vals = (vals.view(rtype).reshape(-1, 2) if np.iscomplexobj(vals)
                    else np.ascontiguousarray(vals, dtype=rtype).reshape(-1, 2))
            g = np.ascontiguousarray(gcf)
            g = g.astype(ctype) if np.iscomplexobj(g) else np.ascontiguousarray(g, dtype=rtype)
            if g.size * g.itemsize != 64 * S * S * 2 * np.dtype(rtype).itemsize:
                raise MigridError(-1, f"gcf must hold 64*S*S complex values (S={S})")
            check(lib().migrid_server_submit(self.h, job_id, N, S, g.ctypes.data, uv.ctypes.data,
                                             vals.ctypes.data, uv.shape[0], SUBMIT_HOST | extra,
                                             ctypes.byref(t)))
        else:
            for x in (uv, vals, gcf):
                if not x.is_cuda or not x.is_contiguous():
                    raise MigridError(-1, "device tensors must be contiguous and on the GPU")
            check(lib().migrid_server_submit(self.h, job_id, N, S, gcf.data_ptr(), uv.data_ptr(),
                                             vals.data_ptr(), uv.shape[0], SUBMIT_DEVICE | extra,
                                             ctypes.byref(t)))
            self._keep[t.value] = (uv, vals, gcf)
        self._N[job_id] = N
        return t.value

    def poll(self, ticket):
        d, ms = ctypes.c_int32(), ctypes.c_double()
        check(lib().migrid_server_poll(self.h, ticket, ctypes.byref(d), ctypes.byref(ms)))
        if d.value:
            self._keep.pop(ticket, None)
        return bool(d.value), ms.value

    def wait(self, ticket):
        ms = ctypes.c_double()
        rc = lib().migrid_server_wait(self.h, ticket, ctypes.byref(ms))
        self._keep.pop(ticket, None)
        check(rc)
        return ms.value

    def read_grid(self, job_id):
        N = self._N[job_id]
        out = np.empty((N, N), np.complex128 if self.cfg.out_type == F64 else np.complex64)
        check(lib().migrid_server_read_grid(self.h, job_id, out.ctypes.data))
        return out

    def drop_job(self, job_id):
        check(lib().migrid_server_drop_job(self.h, job_id))

    def stats(self):
        st = ServerStats()
        check(lib().migrid_server_get_stats(self.h, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            check(lib().migrid_server_destroy(self.h))
            self.h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
