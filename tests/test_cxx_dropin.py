"""The header-only C++ drop-in (include/migrid.hpp) compiles against the C
ABI with a plain host compiler, prints-and-continues on errors like the
reference's CUDA_CHECK_ERR, and (on a GPU) reproduces gridCPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "ska-gpu-grid_amd", "migrid")


def build_dropin(force=True, tiled=False):
    exe = os.path.join(ROOT, "tests", "cxx", "dropin_main_tiled" if tiled else "dropin_main")
    if not force and os.path.exists(exe):
        return exe
    flags = ["-DMIGRID_USE_TILED"] if tiled else []
    subprocess.run(["g++", "-std=c++14", "-O2", "-Wall", *flags, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cxx", "dropin_main.cpp"), "-L", LIBDIR, "-lmigrid",
                    "-Wl,-rpath,$ORIGIN/../../ska-gpu-grid_amd/migrid", "-o", exe], check=True)
    return exe


def test_shim_compiles_and_reports_errors():
    exe = build_dropin()
    r = subprocess.run([exe, "--bad"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "Error -1 in gridGPU" in r.stdout and "gcf_dim=7" in r.stdout


def test_shim_tiled_flag_compiles():
    exe = build_dropin(tiled=True)
    r = subprocess.run([exe, "--bad"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Error -1 in gridGPU" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("tiled", [False, True], ids=["reference_flags", "MIGRID_USE_TILED"])
def test_dropin_program_matches_gridcpu(tmp_path, tiled):
    import migrid
    import oracle as O
    exe = build_dropin(force=False, tiled=tiled)  # built by __graft_entry__.build()
    N, S, npts = 512, 32, 30000
    out = tmp_path / "grid.bin"
    subprocess.run([exe, str(N), str(S), str(npts), str(out)], check=True, timeout=120)
    padded = np.fromfile(out, np.float64).reshape(-1, 2)
    uv, vals = O.ref_inputs(npts, N)
    ref = O.grid_zz(uv, vals, N, O.init_gcf(S), S)
    g = migrid.unpad(padded, N, S)
    assert O.rel_errors(g, ref)[0] < 1e-12
