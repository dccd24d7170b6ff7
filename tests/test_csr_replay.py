"""CPU replay of the CSR-offset kernels (k_chunk_starts, k_lower_bounds in
csrc/k_bin.hip) with a small chunk: offsets[j] = number of sorted keys < j,
checked against numpy's searchsorted, through both the per-chunk scan and the
per-key binary search."""
import numpy as np
import pytest


def chunk_starts(sk, K, C):
    nchunk = (K + C - 1) // C
    start = np.full(nchunk + 1, sk.size, dtype=np.int64)
    for p in range(sk.size):
        hi = sk[p] // C
        lo = 0 if p == 0 else sk[p - 1] // C + 1
        for c in range(lo, hi + 1):
            start[c] = p
    return start


def lower_bounds(sk, K, C, scan_max):
    start = chunk_starts(sk, K, C)
    offs = np.empty(K, dtype=np.int64)
    for c in range(start.size - 1):
        j0, i0, i1 = c * C, start[c], start[c + 1]
        if i1 - i0 > scan_max:
            for e in range(min(C, K - j0)):
                lo, hi = i0, i1
                while lo < hi:
                    mid = lo + (hi - lo) // 2
                    if sk[mid] < j0 + e:
                        lo = mid + 1
                    else:
                        hi = mid
                offs[j0 + e] = lo
            continue
        first = np.full(C, i1, dtype=np.int64)
        for p in range(i0, i1):
            assert j0 <= sk[p] < j0 + C  # the chunk's positions hold its keys only
            if p == 0 or sk[p - 1] != sk[p]:
                first[sk[p] - j0] = p
        run = np.minimum.accumulate(first[::-1])[::-1]
        n = min(C, K - j0)
        offs[j0:j0 + n] = run[:n]
    return offs


@pytest.mark.parametrize("seed,E,K,C,scan_max", [
    (0, 500, 1000, 16, 10 ** 9),   # scan only
    (1, 500, 1000, 16, 4),         # mostly binary search
    (2, 300, 5000, 64, 32),        # sparse keys, long gaps
    (3, 1, 37, 8, 16),             # one element
    (4, 2000, 60, 8, 100),         # many duplicates per key
])
def test_lower_bounds_replay(seed, E, K, C, scan_max):
    rng = np.random.default_rng(seed)
    sk = np.sort(rng.integers(0, K, E))
    np.testing.assert_array_equal(lower_bounds(sk, K, C, scan_max),
                                  np.searchsorted(sk, np.arange(K), side="left"))


def test_lower_bounds_dense_key():
    """one key holding most elements, with keys at both ends of the space"""
    sk = np.sort(np.concatenate([np.full(300, 41), [0, 1, 99, 99, 63]]))
    for scan_max in (10 ** 9, 50):
        np.testing.assert_array_equal(lower_bounds(sk, 100, 8, scan_max),
                                      np.searchsorted(sk, np.arange(100), side="left"))
