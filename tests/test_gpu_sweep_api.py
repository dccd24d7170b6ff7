"""Seeded random sweeps of the other entry points against the oracle.

- Batches of random jobs (grid edges 3..500, supports 2..64, 0..4000
  visibilities) in one persistent launch, two of them sharing one grid with
  different supports; direct runs and graph replays; FP64 within 1e-12 and
  FP32 tiles within 1e-5.
- The drop-in host calls gridGPU<double2,double2> / <double2,float2> /
  <float2,float2> on random shapes with supports up to 128 (scatter) or 64
  (tiled): the padded buffer's grid against the oracle (1e-12 / 2e-6 /
  2e-5), the padding left zero.
- Gridding with a GCF table per visibility, indexes drawn from 0..ngcf-1,
  against the oracle gridding each group with its table (1e-12); an index
  out of range is an error.
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


def cplx_np(p):
    return p[0] + 1j * p[1]


def cplx_t(g):
    g = g.cpu().numpy().astype(np.float64)
    return g[..., 0] + 1j * g[..., 1]


def check(got, want, tol, what):
    m = np.abs(want).max(initial=0.0)
    if m == 0.0:
        assert np.abs(got).max(initial=0.0) == 0.0, what
        return
    err = np.abs(got - want).max() / m
    assert err <= tol, f"{what}: rel max {err:.3e} > {tol}"


def rand_vis(rng, N, S, n):
    if rng.random() < 0.3 and n > 0:
        uv = rng.uniform(0, N, size=2) + rng.normal(scale=2.0, size=(n, 2))
    else:
        uv = rng.uniform(-S, N + S, size=(n, 2))
    return uv.astype(np.float32).astype(np.float64), rng.normal(size=(n, 2))


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_batch_sweep(seed):
    t = torch()
    rng = np.random.default_rng(300 + seed)
    arith, tol = (migrid.ARITH_NATIVE, 1e-12) if seed % 2 == 0 else (migrid.ARITH_F32, 1e-5)
    print(f"[batch sweep] seed {seed} arith {arith}", flush=True)
    jobs, wants = [], []
    for i in range(6):
        N = int(rng.choice([3, 17, 64, 129, int(rng.integers(20, 500))]))
        S = int(2 * rng.integers(1, 33))
        n = int(rng.choice([0, 1, 500, 4000]))
        uv, vals = rand_vis(rng, N, S, n)
        gcf = O.config_gcf(S)
        wants.append(cplx_np(O.grid_zz(uv, vals, N, gcf, S)))
        jobs.append(dict(grid=t.empty((N, N, 2), dtype=t.float64, device="cuda"), uv=dev(uv),
                         vals=dev(vals), gcf=dev(gcf), N=N, S=S))
    # two more jobs into one shared grid, with different supports
    N = int(rng.integers(40, 300))
    shared = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    want_shared = np.zeros((N, N), complex)
    for S in (int(2 * rng.integers(1, 33)), int(2 * rng.integers(1, 33))):
        uv, vals = rand_vis(rng, N, S, 2000)
        gcf = O.config_gcf(S)
        want_shared += cplx_np(O.grid_zz(uv, vals, N, gcf, S))
        jobs.append(dict(grid=shared, uv=dev(uv), vals=dev(vals), gcf=dev(gcf), N=N, S=S))
    ctx = migrid.Context()
    b = migrid.Batch(ctx, jobs, arith=arith)
    stream = t.cuda.Stream()
    for mode in ("run", "graph", "graph"):
        for j in jobs:
            j["grid"].fill_(123.0)  # fresh grids must not depend on what was there
        t.cuda.synchronize()
        if mode == "run":
            b.run(stream)
        else:
            b.graph_launch(stream)
        t.cuda.synchronize()
        for i, (j, w) in enumerate(zip(jobs[:6], wants)):
            check(cplx_t(j["grid"]), w, tol, f"{mode} job {i} N={j['N']} S={j['S']}")
        check(cplx_t(shared), want_shared, tol, f"{mode} shared grid")
    b.close()


@pytest.mark.parametrize("i", range(8))
def test_dropin_sweep(i):
    rng = np.random.default_rng(500 + i)
    pair = ["zz", "zc", "cc"][i % 3]
    tol = {"zz": 1e-12, "zc": 2e-6, "cc": 2e-5}[pair]
    strategy = "scatter" if i % 2 == 0 else "tiled"
    N = int(rng.choice([5, 33, 100, int(rng.integers(8, 400))]))
    S = int(2 * rng.integers(1, 65 if strategy == "scatter" else 33))
    n = int(rng.choice([0, 1, 300, 3000]))
    print(f"[dropin sweep] {i}: {pair} {strategy} N={N} S={S} n={n}", flush=True)
    uv, vals = rand_vis(rng, N, S, n)
    gcf = O.config_gcf(S)
    oracle = {"zz": O.grid_zz, "zc": O.grid_zc, "cc": O.grid_cc}[pair]
    want = cplx_np(oracle(uv, vals, N, gcf, S))
    padded = migrid.gridGPU(uv, vals, gcf, N, S, pair=pair, strategy=strategy)
    check(migrid.unpad(padded, N, S).astype(complex), want, tol, f"{pair} {strategy}")
    base = migrid.grid_base(N, S)
    flat = padded.reshape(-1, 2)
    assert not np.any(flat[:base]) and not np.any(flat[base + N * N:]), "padding written"


@pytest.mark.parametrize("i", range(4))
def test_indexed_sweep(i):
    t = torch()
    rng = np.random.default_rng(700 + i)
    N = int(rng.integers(16, 300))
    S = int(2 * rng.integers(1, 33))
    ngcf = int(rng.integers(1, 6))
    n = int(rng.choice([1, 1000, 5000]))
    print(f"[indexed sweep] {i}: N={N} S={S} ngcf={ngcf} n={n}", flush=True)
    uv, vals = rand_vis(rng, N, S, n)
    idx = rng.integers(0, ngcf, size=n).astype(np.int32)
    tables = [O.config_gcf(S, beta=0.02 * (g + 1)) for g in range(ngcf)]
    want = np.zeros((N, N), complex)
    for g in range(ngcf):
        m = idx == g
        if m.any():
            want += cplx_np(O.grid_zz(uv[m], vals[m], N, tables[g], S))
    grid = t.empty((N, N, 2), dtype=t.float64, device="cuda")
    ctx = migrid.Context()
    ctx.grid_indexed(grid, dev(uv), dev(vals), dev(np.concatenate(tables)), dev(idx), N, S)
    t.cuda.synchronize()
    check(cplx_t(grid), want, 1e-12, "indexed")
    bad = idx.copy()
    bad[rng.integers(0, n)] = ngcf if i % 2 else -1
    with pytest.raises(migrid.MigridError):
        ctx.grid_indexed(grid, dev(uv), dev(vals), dev(np.concatenate(tables)), dev(bad), N, S)
