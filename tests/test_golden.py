"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture bit for bit, and the fixture inputs
are what the reference's own input recipes give (glibc rand stream,
init_gcf).  GPU: every strategy reproduces the fixtures within 1e-12
relative max-norm (FP64), coverage counts bit-exact.
"""
import glob
import math
import os

import numpy as np
import pytest

import migrid
import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRID_CASES = ["impulse_S8_N40", "refstream_S8_N64", "config_S16_N96"]
STRATS = ["scatter", "tiled", "gather", "window"]


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def cplx(p):
    return p[0] + 1j * p[1]


def test_fixture_set_is_complete():
    names = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz")))
    assert names == sorted(GRID_CASES + ["coverage_S8_N64", "compat_S16_N64"])


@pytest.mark.parametrize("name", GRID_CASES)
def test_oracle_reproduces_fixture_bitwise(name):
    f = load(name)
    g = O.grid_zz(f["uv"], f["vals"], int(f["N"]), f["gcf"], int(f["S"]))
    assert np.array_equal(g, f["expected"])


def test_oracle_coverage_and_compat_fixtures():
    f = load("coverage_S8_N64")
    assert np.array_equal(O.grid_zz(f["uv"], f["vals"], 64, f["gcf"], 8, debug1=True), f["expected"])
    assert (f["expected"][0] == np.round(f["expected"][0])).all()
    c = load("compat_S16_N64")
    r2, i2 = O.compute_gcf_phase(c["uv"], float(c["T"]), float(c["w"]))
    assert np.array_equal(r2, c["r2"]) and np.array_equal(i2, c["i2"])
    g = O.grid_const_tap(c["uv"], c["vals"], 64, 16, r2, i2)
    assert np.array_equal(g, c["expected"])


def test_fixture_inputs_follow_the_reference_recipes():
    f = load("refstream_S8_N64")
    uv, vals = O.ref_inputs(256, 64)
    assert np.array_equal(uv, f["uv"]) and np.array_equal(vals, f["vals"])
    assert np.array_equal(O.init_gcf(8), f["gcf"])
    # windows cross every edge: taps off the grid are dropped, not wrapped
    assert f["uv"].min() < 3 and f["uv"].max() > 61


def test_impulse_fixture_is_the_addressed_subkernels():
    """Known answer: the grid is the sum over k of value_k x sub-kernel k
    placed with its tap (0, 0) at (floor(u) - S/2 + 1, floor(v) - S/2 + 1)."""
    f = load("impulse_S8_N40")
    S, N = 8, 40
    G = f["gcf"].reshape(64, S, S, 2)
    want = np.zeros((N, N), complex)
    for k in range(64):
        u, v = f["uv"][k]
        assert 8 * math.floor(8 * (v % 1)) + math.floor(8 * (u % 1)) == k
        x0, y0 = math.floor(u) - S // 2 + 1, math.floor(v) - S // 2 + 1
        vr, vi = f["vals"][k]
        want[y0:y0 + S, x0:x0 + S] += (vr + 1j * vi) * (G[k, :, :, 0] + 1j * G[k, :, :, 1])
    assert np.abs(cplx(f["expected"]) - want).max() < 1e-13


# ---- GPU -------------------------------------------------------------------------

def gpu_grid(f, strategy, **kw):
    import torch
    N, S = int(f["N"]), int(f["S"])
    planar = strategy == "window"
    grid = torch.empty((2, N, N) if planar else (N, N, 2), dtype=torch.float64, device="cuda")
    d = [torch.from_numpy(np.ascontiguousarray(f[k])).cuda() for k in ("uv", "vals", "gcf")]
    migrid.Context().grid(grid, *d, N, S, strategy=strategy, **kw)
    torch.cuda.synchronize()
    g = grid.cpu().numpy()
    return g if planar else np.moveaxis(g, -1, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("name", GRID_CASES)
def test_gpu_reproduces_fixture(name, strategy):
    f = load(name)
    g = gpu_grid(f, strategy)
    assert O.rel_errors(g, f["expected"])[0] < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("strategy", STRATS)
def test_gpu_coverage_fixture_bit_exact(strategy):
    f = load("coverage_S8_N64")
    g = gpu_grid(f, strategy, debug_coverage=True)
    assert np.array_equal(g, f["expected"])


@pytest.mark.gpu
@pytest.mark.parametrize("strategy", ["scatter", "tiled"])
def test_gpu_compat_fixture(strategy):
    c = load("compat_S16_N64")
    S, e = 16, 16 * 15 // 2 - 1
    gcf = np.zeros((64 * S * S, 2))
    gcf[e] = (c["T"], c["w"])
    g = gpu_grid(dict(c, gcf=gcf), strategy, gcf_mode=migrid.GCF_COMPAT)
    # float phase: 1 ulp of sqrtf or of the sin/cos argument moves a tap by
    # about |phase| * 2^-24
    f32 = np.float32
    p1 = f32(2 * 3.1415926 * c["w"])
    p2 = f32(float(p1) * c["T"])
    m = np.floor(c["uv"])
    d = (m - c["uv"] + np.floor(8 * (c["uv"] - m)) / 8.0).astype(f32)
    phase = np.abs(p1 - p2 * np.sqrt((d * d).sum(1)))
    tol = max(1e-6, 8 * float(phase.max()) * 2.0 ** -24)
    assert O.rel_errors(g, c["expected"])[0] < tol


@pytest.mark.gpu
def test_gpu_drop_in_padded_buffer_reproduces_fixture():
    f = load("refstream_S8_N64")
    padded = migrid.gridGPU(f["uv"], f["vals"], f["gcf"], 64, 8)
    assert O.rel_errors(migrid.unpad(padded, 64, 8), cplx(f["expected"]))[0] < 1e-12
    # the padding around the grid stays zero
    base = migrid.grid_base(64, 8)
    assert not padded[:base].any() and not padded[base + 64 * 64:].any()
