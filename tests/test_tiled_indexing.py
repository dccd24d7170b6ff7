"""CPU replay of the tiled owner kernel's index arithmetic.

For real bin sorts (the oracle's, bit-identical to the device sort) this
walks every (tile, wave, sub-kernel, bin row) the kernel visits and checks
that every CSR read is inside the offsets array, every record range is a
valid slice of the records, and every staged-LDS element a lane reads lies in
the padded buffer without wrapping a row.  It is the regression test for
edge tiles, whose waves past the grid's last bin must see empty ranges.
"""
import numpy as np
import pytest

import oracle as O

H = 16


def border(S):
    return 16 * ((S + 30) // 16) - S


def rows(S):
    return S + 2 * border(S)


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
                    for k in range(64):
                        for by in range(byl, byh + 1):
                            key = (k * nb + by) * nb
                            i0, i1 = key + bxl, key + bxh + 1
                            assert 0 <= i0 <= nkeys and 0 <= i1 <= nkeys
                            r0, r1 = int(offs[i0]), int(offs[i1])
                            assert 0 <= r0 <= r1 <= count
                            if r0 == r1:
                                continue
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
