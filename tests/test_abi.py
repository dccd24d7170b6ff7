"""CPU checks of the C ABI: the library loads, exports every symbol the
header declares, and rejects bad arguments before touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import migrid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:[A-Za-z_][\w \*]*?)\b(migrid_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = migrid.lib()
    decl = declared_symbols()
    assert "migrid_gridGPU_zz" in decl and len(decl) >= 15
    for name in sorted(decl):
        assert hasattr(L, name), name
    assert set(migrid.EXPORTS) <= decl


def test_version_and_defaults():
    assert "gfx950" in migrid.version()
    o = migrid.make_opts()
    assert o.strategy == migrid.SCATTER and o.layout == 0 and o.accumulate == 0
    assert ctypes.sizeof(migrid.Opts) == 4 * 8 + 8 * 2 + 4 * 8
    assert migrid.padded_elems(2048, 32) == 2048 * 2048 + 2 * 2048 * 32 + 2 * 32


@pytest.mark.parametrize("N,S,npts", [(64, 7, 10), (64, 0, 10), (64, 130, 10), (0, 8, 10), (64, 8, -1)])
def test_bad_shapes_rejected_before_any_device_call(N, S, npts):
    L = migrid.lib()
    buf = np.zeros(16, np.float64)
    rc = L.migrid_gridGPU_zz(buf.ctypes.data, buf.ctypes.data, buf.ctypes.data,
                             ctypes.c_size_t(npts & 0xFFFFFFFFFFFFFFFF), N, buf.ctypes.data, S, None)
    assert rc == -1
    assert L.migrid_last_error().decode()


def test_null_pointers_rejected():
    L = migrid.lib()
    buf = np.zeros(16, np.float64)
    assert L.migrid_gridGPU_zz(None, buf.ctypes.data, buf.ctypes.data, 1, 64, buf.ctypes.data, 8, None) == -1
    assert "out" in L.migrid_last_error().decode()
    assert L.migrid_gridGPU_cc(buf.ctypes.data, None, None, 3, 64, buf.ctypes.data, 8, None) == -1
    assert L.migrid_gridGPU_zc(buf.ctypes.data, buf.ctypes.data, buf.ctypes.data, 1, 64, None, 8, None) == -1
    assert L.migrid_ctx_grid(None, None, None, None) == -1
    assert L.migrid_grid_noalloc(None, None, None) == -1
    assert "job" in L.migrid_last_error().decode()


def test_gcf_size_checked_by_binding():
    with pytest.raises(migrid.MigridError):
        migrid.gridGPU(np.zeros((1, 2)), np.zeros((1, 2)), np.zeros((10, 2)), 64, 8)


def test_bins_per_axis():
    L = migrid.lib()
    assert L.migrid_bins_per_axis(8192, 64, 64) == (8192 + 62) // 64 + 1
    assert L.migrid_bins_per_axis(100, 8, 0) == 0


def test_make_opts_keeps_every_field():
    o = migrid.make_opts(strategy="tiled", tile=32, vis_per_wave=777, stage_min=5, wg_per_cu=2,
                         accumulate=True, sigma=3.0, beta=0.1, tap_smax=-1)
    assert (o.strategy, o.tile, o.vis_per_wave, o.stage_min, o.wg_per_cu, o.accumulate,
            o.tap_smax) == (migrid.TILED, 32, 777, 5, 2, 1, -1)
    assert o.gcf_sigma == 3.0 and o.gcf_beta == 0.1


def test_server_structs_and_defaults():
    assert ctypes.sizeof(migrid.ServerCfg) == 6 * 4 + 8 + ctypes.sizeof(migrid.Opts)
    assert ctypes.sizeof(migrid.ServerStats) == 14 * 8  # 11 int64 + 3 double
    c = migrid.ServerCfg()
    migrid.lib().migrid_server_cfg_default(ctypes.byref(c))
    assert (c.in_type, c.out_type, c.max_batch, c.max_inflight) == (migrid.F64, migrid.F64, 64, 2)
    assert c.opts.strategy == migrid.TILED and c.opts.arith == migrid.ARITH_F32


@pytest.mark.parametrize("field,value", [("in_type", 7), ("out_type", -1), ("max_inflight", 9),
                                         ("max_batch", -2), ("hbm_budget", -5)])
def test_server_config_rejected_before_any_device_call(field, value):
    L = migrid.lib()
    c = migrid.ServerCfg()
    L.migrid_server_cfg_default(ctypes.byref(c))
    setattr(c, field, value)
    h = ctypes.c_void_p()
    assert L.migrid_server_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert not h.value and L.migrid_last_error().decode()


def test_server_null_handles_rejected():
    L = migrid.lib()
    t = ctypes.c_int64()
    assert L.migrid_server_submit(None, 1, 64, 8, None, None, None, 0, 0, ctypes.byref(t)) == -1
    assert L.migrid_server_wait(None, 1, None) == -1
    assert L.migrid_server_read_grid(None, 1, None) == -1
    assert L.migrid_server_destroy(None) == 0


def test_new_entry_points_reject_bad_arguments_before_any_device_call():
    L = migrid.lib()
    d = (ctypes.c_double * 8)()
    i = (ctypes.c_int32 * 4)()
    # NULL contexts
    assert L.migrid_ctx_fft(None, migrid.F64, migrid.LAYOUT_INTERLEAVED, d, d, 8,
                            migrid.FFT_TO_IMAGE, 1.0, None, None) == -1
    j = migrid.Job()
    assert L.migrid_ctx_degrid(None, ctypes.byref(j), d, None, None) == -1
    assert L.migrid_ctx_grid_indexed(None, ctypes.byref(j), 1, i, None, None) == -1
    # uv -> pixel conversion: shapes and parameters checked first
    assert L.migrid_uvw_to_pixels(d, 1, d, 0, 64, 1e-4, 0, 1.0, d, None, None) == -1   # nchan
    assert L.migrid_uvw_to_pixels(d, 1, d, 1, 64, 0.0, 0, 1.0, d, None, None) == -1    # cell
    assert L.migrid_uvw_to_pixels(d, 1, d, 1, 64, 1e-4, 4, 0.0, d, i, None) == -1      # w_max
    assert L.migrid_uvw_to_pixels(d, 1, d, 1, 64, 1e-4, 4, 1.0, d, None, None) == -1   # no index out
    assert L.migrid_uvw_to_pixels(None, 1, d, 1, 64, 1e-4, 0, 1.0, d, None, None) == -1  # NULL uvw
    assert L.migrid_uvw_to_pixels(None, 0, None, 1, 64, 1e-4, 0, 1.0, None, None, None) == 0  # empty
    # the text reader checks its arguments
    n = ctypes.c_int64()
    assert L.migrid_read_uvw_text(None, 64, 1, d, None, ctypes.byref(n)) == -1
    assert L.migrid_read_uvw_text(b"/nonexistent/uvw.dat", 64, 1, d, None, ctypes.byref(n)) == -1
    assert "cannot open" in L.migrid_last_error().decode()
    assert L.migrid_read_uvw_text(b"x", 0, 1, d, None, ctypes.byref(n)) == -1  # img_dim
