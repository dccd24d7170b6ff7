// Shared types of the tiled owner gridder (k_tiled.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace migrid {

// Packed visibility records in bin order.  q = q_x | q_y << 16 with
// q = floor(coordinate) + S/2 (so every binned visibility has 0 <= q < 2^16).
struct alignas(16) RecF32 {
    uint32_t q;
    uint32_t n;
    float vr, vi;
};
struct alignas(16) RecF64 {
    uint32_t q;
    uint32_t n;
    uint32_t pad0, pad1;
    double vr, vi;
};

// One unit of work: a tile and an inclusive sub-kernel range.  split = the
// tile has more than one item, so its flush must be atomic.
struct Item {
    int32_t tile;
    int16_t klo, khi;
};

struct QueueState {
    int32_t nitems;
    int32_t head;
    int32_t pad[14];
};

struct TileGeom {
    int N, S;
    int T;    // tile edge (pixels)
    int H;    // bin edge (pixels)
    int nb;   // bins per axis
    int ntx;  // tiles per axis
};

}  // namespace migrid
