"""CPU replay of the tiled owner kernel's index arithmetic.

For real bin sorts (the oracle's, bit-identical to the device sort) this
walks every (tile, wave, sub-kernel, bin row) the kernel visits and checks
that every CSR read is inside the offsets array, every record range is a
valid slice of the records, and every staged-LDS element a lane reads lies in
the padded buffer without wrapping a row.  It is the regression test for
edge tiles, whose waves past the grid's last bin must see empty ranges.

A second replay runs the work-item pass (k_items) on the same sorts and
checks that the partial-tile slots and reduce entries the plan allocates for
the ordered reduction of gather / window cover every split tile.
"""
import numpy as np
import pytest

import oracle as O

H = 16


def border(S):
    return 16 * ((S + 30) // 16) - S


def rows(S):
    # the zero border before the kernel covers the lowest tap a lane reads;
    # 15 rows / columns after it the highest (lane offset up to 15)
    return S + border(S) + 15


def pitch(S):
    r = rows(S)
    return r + ((8 - r % 32) + 32) % 32


@pytest.mark.parametrize("N,S,T", [(70, 6, 64), (100, 8, 32), (130, 64, 64), (48, 16, 16),
                                   (257, 32, 64), (64, 64, 32), (33, 2, 16)])
def test_owner_indexing_stays_in_bounds(N, S, T):
    uv, _ = O.gen_uniform(3000, -S, N + S, seed=N * 7 + S)
    perm, offs = O.bin_sort(uv, N, S, H)
    nb = O.n_bins(N, S, H)
    nkeys = 64 * nb * nb
    assert offs.size == nkeys + 1
    count = int(offs[-1])
    m = np.floor(uv[perm]).astype(np.int64)
    q = m + S // 2
    B, P, R = border(S), pitch(S), rows(S)
    ntx = (N + T - 1) // T
    NW1 = T // 16
    for ty in range(ntx):
        for tx in range(ntx):
            X0, Y0 = tx * T, ty * T
            for wy in range(NW1):
                for wx in range(NW1):
                    Xw, Yw = X0 + 16 * wx, Y0 + 16 * wy
                    bxl, bxh = Xw // H, min(nb - 1, (Xw + S + 14) // H)
                    byl, byh = Yw // H, min(nb - 1, (Yw + S + 14) // H)
                    wave_on = Xw < N and Yw < N and bxl <= bxh and byl <= byh
                    if not wave_on:
                        continue
                    assert byh - byl + 1 <= 5
                    # the visibilities whose window reaches the block's pixels on
                    # the grid: each must be listed for this wave exactly once
                    xe, ye = min(Xw + 15, N - 1), min(Yw + 15, N - 1)
                    reach = ((q[:, 0] >= Xw) & (q[:, 0] - S + 1 <= xe) &
                             (q[:, 1] >= Yw) & (q[:, 1] - S + 1 <= ye))
                    listed = np.zeros(count, np.int64)
                    for k in range(64):
                        for by in range(byl, byh + 1):
                            key = (k * nb + by) * nb
                            i0, i1 = key + bxl, key + bxh + 1
                            assert 0 <= i0 <= nkeys and 0 <= i1 <= nkeys
                            r0, r1 = int(offs[i0]), int(offs[i1])
                            assert 0 <= r0 <= r1 <= count
                            if r0 == r1:
                                continue
                            listed[r0:r1] += 1
                            dx = q[r0:r1, 0] - Xw
                            dy = q[r0:r1, 1] - Yw
                            assert dx.min() >= 0 and dy.min() >= 0
                            # element column/row read by lanes lx + 8 j in [0, 15]
                            col_lo = (S - 1 + B) - dx.max()
                            col_hi = (S - 1 + B) - dx.min() + 15
                            row_lo = (S - 1 + B) - dy.max()
                            row_hi = (S - 1 + B) - dy.min() + 15
                            assert col_lo >= 0 and col_hi < R and col_hi < P
                            assert row_lo >= 0 and row_hi < R
                            assert row_hi * P + col_hi < R * P
                    assert (listed[reach] == 1).all() and listed.max() <= 1


def tile_counts(offs, N, S, T):
    """tcount[tile, k] as k_tile_counts computes it: the visibilities of
    sub-kernel k in the bins tile (tx, ty) lists, bins tx*T/H ..
    min(nb - 1, (tx*T + T + S - 2) / H) per axis."""
    nb = O.n_bins(N, S, H)
    ntx = (N + T - 1) // T
    # 2-D prefix sums over (by, bx) of the per-key counts
    cnt = np.diff(offs.astype(np.int64))[: 64 * nb * nb].reshape(64, nb, nb)
    P = np.zeros((64, nb + 1, nb + 1), np.int64)
    P[:, 1:, 1:] = cnt.cumsum(1).cumsum(2)
    t = np.arange(ntx)
    lo = t * T // H
    hi = np.minimum(nb - 1, (t * T + T + S - 2) // H) + 1
    tc = (P[:, hi[:, None], hi[None, :]] - P[:, lo[:, None], hi[None, :]]
          - P[:, hi[:, None], lo[None, :]] + P[:, lo[:, None], lo[None, :]])
    return tc.reshape(64, ntx * ntx).T  # [tile, k]


def items_per_tile(tc, target):
    """The number of work items k_items cuts each tile into."""
    total = tc.sum(1)
    nit = np.minimum(64, -(-total // target))
    inc = tc.cumsum(1)
    excl = inc - tc
    with np.errstate(divide="ignore", invalid="ignore"):
        j = np.where(total[:, None] > 0, excl * nit[:, None] // np.maximum(total, 1)[:, None], 0)
    n = np.zeros(tc.shape[0], np.int64)
    for t in np.nonzero(total)[0]:
        nz = np.nonzero(tc[t])[0]
        jj = j[t, nz]
        n[t] = 1 + int((jj[1:] != jj[:-1]).sum())
    return n


def plan_caps(E, N, S, T, target):
    """The ordered reduction's capacities as the plan sizes them."""
    per_axis = (T + S - 2) // T + 1
    m = per_axis * per_axis
    return 2 * m * E // target + 1, m * E // target + 1, m


@pytest.mark.parametrize("N,S,T,npts,mean_r,target", [
    (1024, 64, 64, 300_000, 60.0, 4096), (1024, 64, 64, 300_000, 60.0, 256),
    (512, 32, 32, 200_000, 30.0, 512), (512, 16, 16, 100_000, 20.0, 128),
    (700, 64, 16, 100_000, 40.0, 1024), (8192, 64, 64, 10_000_000, 800.0, 4096)])
def test_ordered_reduction_slots_cover_every_split_tile(N, S, T, npts, mean_r, target):
    uv, _ = O.gen_clustered(npts, N, S, mean_r=mean_r, seed=1541618)
    perm, offs = O.bin_sort(uv, N, S, H)
    E = perm.size
    tc = tile_counts(offs, N, S, T)
    pcap, rcap, m = plan_caps(E, N, S, T, target)
    assert tc.sum() <= m * E
    n = items_per_tile(tc, target)
    split = n > 1
    need = int(n[split].sum())
    print(f"  sum(tcount) = {tc.sum() / E:.2f} E, {int(split.sum())} split tiles, "
          f"{need} slots (cap {pcap}, reduce cap {rcap})")
    assert need <= pcap
    assert int(split.sum()) <= rcap
