// Bin-sort geometry and entry points (k_bin.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace migrid {

// Bins of H pixels over q = floor(coordinate) + S/2, q in [0, N+S-2].
// A pixel x receives contributions from visibilities with q in [x, x+S-1].
struct BinGeom {
    int N = 0, S = 0, H = 0;
    int nb = 0;         // bins per axis
    int64_t nkeys = 0;  // 64 * nb * nb
};

BinGeom bin_geom(int64_t N, int64_t S, int H);
size_t bin_scratch_bytes(int64_t npts, const BinGeom& g);

// perm[0..count): stable key order of binned visibility indices;
// offsets[0..nkeys]: CSR starts, offsets[nkeys] = count.  Asynchronous.
int bin_sort(const BinGeom& g, int in_type, const void* uv, int64_t npts, void* scratch,
             int32_t* perm, int32_t* offsets, hipStream_t s);

int64_t scan_scratch_elems(int64_t n);
int exclusive_scan(const int32_t* in, int64_t n, int32_t* out, int32_t* partial, bool write_total,
                   hipStream_t s);

}  // namespace migrid
