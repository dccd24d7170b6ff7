#!/usr/bin/env python3
"""Writes the golden fixtures in this directory (run from the repo root:
python tests/golden/make_golden.py).

The reference ships no golden vectors, and its gridCPU cannot be built here
(it includes cuda.h).  These fixtures are computed by a numpy brute force
written from the reference's per-visibility contract (grid.cu:87-139:
main = floor, sub = floor(8 frac), k = 8 sub_y + sub_x, tap (r, c) of
sub-kernel k lands on (main_x + c - S/2 + 1, main_y + r - S/2 + 1), taps
off the grid dropped, re = vr gr - vi gi, im = vr gi + vi gr, visibilities
added in index order).  That brute force does not share code with the C
oracle; make_golden.py asserts the oracle reproduces every fixture bit for
bit before writing it.

Fixtures (npz, planar (2, N, N) float64 grids as gridCPU writes them):
  impulse_S8_N40     one complex visibility per sub-kernel (a known answer)
  refstream_S8_N64   the reference's default inputs: glibc srand(1541617) /
                     rand() uv in [0, N) and values in [0, 1] (grid.cu:444-452),
                     its init_gcf table (grid.cu:72-85); windows cross the edges
  config_S16_N96     Gaussian x chirp GCF, uniform uv over the interior,
                     N(0,1) values (the 2048^2 workload's form, small)
  coverage_S8_N64    DEBUG1 coverage counts (grid.cu:127-129): integers
  compat_S16_N64     COMPUTE_GCF taps (grid_gpu.cu:137-147) from the oracle's
                     float phase, gridded by the brute force
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402


def init_gcf_np(S):
    """grid.cu:72-85 evaluated in numpy."""
    g = np.empty((64 * S * S, 2))
    for sx in range(8):
        for sy in range(8):
            for x in range(S):
                for y in range(S):
                    tmp = math.sin(6.28 * x / S / 8) * math.exp(-(1.0 * x * x + 1.0 * y * y * sy) / S / S / 2)
                    i = S * S * (sx + sy * 8) + x + y * S
                    g[i, 0] = tmp * math.sin(1.0 * x * sx / (y + 1)) + 0.4
                    g[i, 1] = tmp * math.cos(1.0 * x * sx / (y + 1)) - 0.2
    return g


def config_gcf_np(S, beta=0.05):
    sigma = S / 6.0
    sy, sx, r, c = np.meshgrid(np.arange(8), np.arange(8), np.arange(S), np.arange(S), indexing="ij")
    dx = (c - S // 2 + 1) - sx / 8.0
    dy = (r - S // 2 + 1) - sy / 8.0
    q = dx * dx + dy * dy
    amp = np.exp(-q / (2 * sigma * sigma))
    return np.stack([amp * np.cos(beta * q), amp * np.sin(beta * q)], -1).reshape(-1, 2)


def brute(uv, vals, N, gcf, S, debug1=False, taps=None):
    """Per-visibility gridding in index order.  taps: optional per-vis
    constant tap values (r2, i2) replacing the table (COMPUTE_GCF)."""
    out = np.zeros((2, N, N))
    G = gcf.reshape(64, S, S, 2) if gcf is not None else None
    h = S // 2
    for n in range(uv.shape[0]):
        u, v = uv[n]
        mx, my = math.floor(u), math.floor(v)
        sx, sy = math.floor(8 * (u - mx)), math.floor(8 * (v - my))
        k = 8 * sy + sx
        x0, y0 = mx - h + 1, my - h + 1
        c0, c1 = max(0, -x0), min(S, N - x0)
        r0, r1 = max(0, -y0), min(S, N - y0)
        if c0 >= c1 or r0 >= r1:
            continue
        if debug1:
            out[0, y0 + r0:y0 + r1, x0 + c0:x0 + c1] += 1.0
            out[1, y0 + r0:y0 + r1, x0 + c0:x0 + c1] += float(n)
            continue
        if taps is not None:
            gr = np.full((r1 - r0, c1 - c0), float(taps[0][n]))
            gi = np.full((r1 - r0, c1 - c0), float(taps[1][n]))
        else:
            gr = G[k, r0:r1, c0:c1, 0]
            gi = G[k, r0:r1, c0:c1, 1]
        vr, vi = vals[n]
        out[0, y0 + r0:y0 + r1, x0 + c0:x0 + c1] += vr * gr - vi * gi
        out[1, y0 + r0:y0 + r1, x0 + c0:x0 + c1] += vr * gi + vi * gr
    return out


def impulse_case():
    S, N = 8, 40
    gcf = init_gcf_np(S)
    uv, vals = [], []
    for k in range(64):
        sx, sy = k % 8, k // 8
        mx, my = 4 + 4 * (k % 8), 4 + 4 * (k // 8)
        uv.append((mx + sx / 8 + 1 / 32, my + sy / 8 + 1 / 64))
        vals.append((1.0 + k / 16, (k % 5) - 2.0))
    return dict(uv=np.array(uv), vals=np.array(vals), gcf=gcf, N=N, S=S)


def refstream_case():
    S, N = 8, 64
    uv, vals = O.ref_inputs(256, N)
    return dict(uv=uv, vals=vals, gcf=init_gcf_np(S), N=N, S=S)


def config_case():
    S, N = 16, 96
    uv, vals = O.gen_uniform(1500, S / 2, N - S / 2, seed=20240)
    return dict(uv=uv, vals=vals, gcf=config_gcf_np(S), N=N, S=S)


def main():
    cases = {"impulse_S8_N40": impulse_case(), "refstream_S8_N64": refstream_case(),
             "config_S16_N96": config_case()}
    for name, c in cases.items():
        exp = brute(c["uv"], c["vals"], c["N"], c["gcf"], c["S"])
        orc = O.grid_zz(c["uv"], c["vals"], c["N"], c["gcf"], c["S"])
        assert np.array_equal(exp, orc), name
        np.savez_compressed(os.path.join(HERE, name + ".npz"), expected=exp, **c)
        print(name, "max |G|", np.abs(exp).max())
    # coverage counts (integer valued)
    c = refstream_case()
    exp = brute(c["uv"], c["vals"], c["N"], None, c["S"], debug1=True)
    orc = O.grid_zz(c["uv"], c["vals"], c["N"], c["gcf"], c["S"], debug1=True)
    assert np.array_equal(exp, orc)
    np.savez_compressed(os.path.join(HERE, "coverage_S8_N64.npz"), expected=exp, **c)
    # COMPUTE_GCF: T, w from the table entry the reference reads (grid_gpu.cu:103-108)
    S, N = 16, 64
    uv, vals = O.gen_uniform(400, S / 2, N - S / 2, seed=777)
    gcf = config_gcf_np(S)
    gcf[S * (S - 1) // 2 - 1] = (0.01, 0.25)
    T, w = gcf[S * (S - 1) // 2 - 1]
    r2, i2 = O.compute_gcf_phase(uv, T, w)
    exp = brute(uv, vals, N, None, S, taps=(r2, i2))
    orc = O.grid_const_tap(uv, vals, N, S, r2, i2)
    assert np.array_equal(exp, orc)
    np.savez_compressed(os.path.join(HERE, "compat_S16_N64.npz"), expected=exp, uv=uv, vals=vals,
                        T=T, w=w, N=N, S=S, r2=r2, i2=i2)
    print("written to", HERE)


if __name__ == "__main__":
    main()
