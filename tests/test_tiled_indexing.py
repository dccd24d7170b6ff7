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


def tile_counts(offs, N, S, T, Ty=None, Hb=H):
    """tcount[tile, k] as k_tile_counts computes it: the visibilities of
    sub-kernel k in the bins tile (tx, ty) lists, bins tx*T/H ..
    min(nb - 1, (tx*T + T + S - 2) / H) along x, the same with Ty along y."""
    Ty = T if Ty is None else Ty
    nb = O.n_bins(N, S, Hb)
    ntx, nty = (N + T - 1) // T, (N + Ty - 1) // Ty
    # 2-D prefix sums over (by, bx) of the per-key counts
    cnt = np.diff(offs.astype(np.int64))[: 64 * nb * nb].reshape(64, nb, nb)
    P = np.zeros((64, nb + 1, nb + 1), np.int64)
    P[:, 1:, 1:] = cnt.cumsum(1).cumsum(2)
    tx, ty = np.arange(ntx), np.arange(nty)
    xlo = tx * T // Hb
    xhi = np.minimum(nb - 1, (tx * T + T + S - 2) // Hb) + 1
    ylo = ty * Ty // Hb
    yhi = np.minimum(nb - 1, (ty * Ty + Ty + S - 2) // Hb) + 1
    tc = (P[:, yhi[:, None], xhi[None, :]] - P[:, ylo[:, None], xhi[None, :]]
          - P[:, yhi[:, None], xlo[None, :]] + P[:, ylo[:, None], xlo[None, :]])
    return tc.reshape(64, nty * ntx).T  # [tile, k]


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


def plan_caps(E, N, S, T, target, Ty=None):
    """The ordered reduction's capacities as the plan sizes them."""
    Ty = T if Ty is None else Ty
    m = ((T + S - 2) // T + 1) * ((Ty + S - 2) // Ty + 1)
    return 2 * m * E // target + 1, m * E // target + 1, m


# tile 16 and 32: square tiles with bins of 16; tile 64: the column kernel's
# 128 x 64 tiles with bins of 8
def tile_geom(T):
    return (128, 64, 8) if T == 64 else (T, T, 16)


@pytest.mark.parametrize("N,S,T,npts,mean_r,target", [
    (1024, 64, 64, 300_000, 60.0, 4096), (1024, 64, 64, 300_000, 60.0, 256),
    (512, 32, 32, 200_000, 30.0, 512), (512, 16, 16, 100_000, 20.0, 128),
    (700, 64, 16, 100_000, 40.0, 1024), (700, 18, 64, 100_000, 40.0, 128),
    (8192, 64, 64, 10_000_000, 800.0, 4096)])
def test_ordered_reduction_slots_cover_every_split_tile(N, S, T, npts, mean_r, target):
    uv, _ = O.gen_clustered(npts, N, S, mean_r=mean_r, seed=1541618)
    Tx, Ty, Hb = tile_geom(T)
    perm, offs = O.bin_sort(uv, N, S, Hb)
    E = perm.size
    tc = tile_counts(offs, N, S, Tx, Ty, Hb)
    pcap, rcap, m = plan_caps(E, N, S, Tx, target, Ty)
    assert tc.sum() <= m * E
    n = items_per_tile(tc, target)
    split = n > 1
    need = int(n[split].sum())
    print(f"  sum(tcount) = {tc.sum() / E:.2f} E, {int(split.sum())} split tiles, "
          f"{need} slots (cap {pcap}, reduce cap {rcap})")
    assert need <= pcap
    assert int(split.sum()) <= rcap


# ---- the column-block owner kernel (128 x 64 tiles, bins of 8) -------------------
CH = 8   # bin edge
CG = 8   # row groups per lane
CTX, CTY = 128, 64  # tile width and height


def col_depth(S):
    return (S + 6) // 8


def col_lead(S):
    return 8 * col_depth(S) + 8 - S


def col_pitch(S):
    p = S + col_lead(S)
    return p if p % 16 == 8 else p + 8


def col_rows(S):
    return S + col_lead(S) + 8


def col_ncols(S):
    return col_depth(S) + 1


def col_nrows(S):
    return (8 * CG + S - 2) // 8 + 1


@pytest.mark.parametrize("S", list(range(2, 66, 2)))
def test_column_layout_constants(S):
    """Pitch 8 or 24 mod 32 (bank-conflict free), a multiple of 8 (row
    groups pair in ds_read2st64), at least S + lead (shared border), and the
    layout of every smaller support fits in the largest one's buffer; the
    largest group offset of 16-byte elements fits ds_read_b128's 16-bit
    immediate, the bounds of a wave's bin rows fit one vector of 64 lanes."""
    L, P, R = col_lead(S), col_pitch(S), col_rows(S)
    assert 8 <= L <= 15
    assert P % 32 in (8, 24) and P % 8 == 0 and P >= S + L
    # the last row a lane reads (S + 6) spills at most into row S + 7 + L
    assert (S + 6 + L) * P + (S + 6 + L) < R * P
    for s2 in range(2, S + 1, 2):
        assert col_rows(s2) * col_pitch(s2) <= R * P
    assert (CG - 1) * 8 * P * 16 < 65536
    assert (CG - 1) * P // 8 < 256  # ds_read2st64 offsets of 8- and 4-byte elements
    assert 2 * col_nrows(S) <= 64
    # every bin row's group range is non-empty
    D = col_depth(S)
    for t in range(col_nrows(S)):
        assert max(0, t - D) <= min(CG - 1, t)


@pytest.mark.parametrize("N,S,npts", [(130, 64, 3000), (257, 32, 3000), (70, 6, 2000),
                                      (100, 18, 2000), (64, 64, 2000), (200, 48, 2500),
                                      (90, 58, 2500), (150, 40, 2500), (300, 64, 4000)])
def test_column_owner_indexing(N, S, npts):
    """Replays k_owner_col's reads for real bin sorts: every CSR index and
    record range is valid, every visibility whose window reaches a wave's
    block is listed for that wave once, every staged element a lane reads
    lies in the LDS buffer and is the visibility's tap when the tap is in the
    window and a zero otherwise, and no row group a bin row skips holds a
    tap of its visibilities."""
    uv, _ = O.gen_uniform(npts, -S, N + S, seed=N * 11 + S)
    perm, offs = O.bin_sort(uv, N, S, CH)
    nb = O.n_bins(N, S, CH)
    nkeys = 64 * nb * nb
    assert offs.size == nkeys + 1
    count = int(offs[-1])
    q = np.floor(uv[perm]).astype(np.int64) + S // 2
    L, P, R = col_lead(S), col_pitch(S), col_rows(S)
    D, NC, NR = col_depth(S), col_ncols(S), col_nrows(S)
    lx = np.tile(np.arange(8), 8)
    ly = np.repeat(np.arange(8), 8)
    ntx, nty = (N + CTX - 1) // CTX, (N + CTY - 1) // CTY
    for ty in range(nty):
        for tx in range(ntx):
            X0, Y0 = CTX * tx, CTY * ty
            for w in range(16):
                Xw, Yw = X0 + 8 * w, Y0
                on = Xw < N and Yw < N
                bxl, byl = Xw // CH, Yw // CH
                ncx = min(nb - bxl, NC) if on else 0
                nry = min(nb - byl, NR) if on and ncx > 0 else 0
                if nry <= 0:
                    continue
                assert 2 * nry <= 64
                xe, ye = min(Xw + 7, N - 1), min(Yw + 8 * CG - 1, N - 1)
                reach = ((q[:, 0] >= Xw) & (q[:, 0] - S + 1 <= xe) &
                         (q[:, 1] >= Yw) & (q[:, 1] - S + 1 <= ye))
                listed = np.zeros(count, np.int64)
                for k in range(64):
                    key0 = (k * nb + byl) * nb + bxl
                    for t in range(nry):
                        i0, i1 = key0 + t * nb, key0 + t * nb + ncx
                        assert 0 <= i0 <= i1 <= nkeys
                        r0, r1 = int(offs[i0]), int(offs[i1])
                        assert 0 <= r0 <= r1 <= count
                        if r0 == r1:
                            continue
                        listed[r0:r1] += 1
                        dx = q[r0:r1, 0] - Xw
                        dy = q[r0:r1, 1] - Yw
                        assert dx.min() >= 0 and dx.max() <= 8 * NC - 1
                        assert dy.min() >= 8 * t and dy.max() <= 8 * t + 7
                        A, B = max(0, t - D), min(CG - 1, t)
                        c = lx[None, :] - dx[:, None] + S - 1          # [vis, lane]
                        for i in range(CG):
                            r = ly[None, :] + 8 * i - dy[:, None] + S - 1
                            tap = (r >= 0) & (r < S) & (c >= 0) & (c < S)
                            pix = (Xw + lx[None, :] < N) & (Yw + ly[None, :] + 8 * i < N)
                            if i < A or i > B:
                                # a skipped group holds no tap on the grid
                                assert not (tap & pix).any()
                                continue
                            idx = (r + L) * P + (c + L)
                            assert idx.min() >= 0 and idx.max() < R * P
                            # data element (rr, cc) sits at (rr + L) P + cc + L
                            rr, cc = idx // P - L, idx % P - L
                            data = (rr >= 0) & (rr < S) & (cc >= 0) & (cc < S)
                            assert (data == tap).all()
                            assert ((rr == r) & (cc == c))[tap].all()
                assert (listed[reach] == 1).all() and listed.max() <= 1
