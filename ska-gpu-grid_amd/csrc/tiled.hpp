// Shared types of the tiled owner gridder (k_tiled.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "bin.hpp"
#include "internal.hpp"

namespace migrid {

// Packed visibility records in bin order.  q = q_x | q_y << 16 with
// q = floor(coordinate) + S/2 (so every binned visibility has 0 <= q < 2^16).
struct alignas(16) RecF32 {
    uint32_t q;
    uint32_t off;  // (q_y * pitch + q_x) x staged element bytes, in the staged-kernel layout
    float vr, vi;
};
struct alignas(16) RecF64 {
    uint32_t q;
    uint32_t off;
    uint32_t pad0, pad1;
    double vr, vi;
};

// One unit of work: a job, a tile of it and an inclusive sub-kernel range.
// pslot >= 0: the tile is split and the item writes its partial tile to
// that slot for the ordered reduction (gather / window), else -1.
struct Item {
    int32_t tile;
    int16_t klo, khi;
    int32_t job;
    int32_t pslot;
};

struct QueueState {
    int32_t nitems;
    int32_t head;
    int32_t npartial;  // partial-tile slots taken
    int32_t nreduce;   // split tiles listed for the ordered reduction
    int32_t overflow;  // split tiles that found no partial slot (none by the plan's bound)
    int32_t pad[11];
};

// A split tile of the ordered-reduction mode: its n partial tiles are slots
// [pslot, pslot + n), summed in that order (n = 0: nothing to reduce).
struct ReduceItem {
    int32_t job, tile, pslot, n;
};

struct TileGeom {
    int N, S;
    int T;    // tile width (pixels)
    int H;    // bin edge (pixels)
    int nb;   // bins per axis
    int ntx;  // tiles along x
    int Ty;   // tile height: T, or 64 for the column kernel's 128 x 64 tiles
    int nty;  // tiles along y
};

// Everything the owner kernel needs about one job (device-resident table,
// indexed by Item::job).  split[tile] != 0: the tile has more than one item,
// so its flush must be atomic.
struct JobDesc {
    TileGeom g;
    const void* rec;
    const int32_t* offs;
    const int32_t* tcount;  // visibilities per (tile, k): tcount[tile * 64 + k]
    const uint8_t* split;
    const void* gcf;
    void* grid;
    void* grid_lo;
    double sigma, beta;
    int32_t fresh;   // grid not cleared beforehand: unsplit tiles are stored,
                     // empty and split tiles were zeroed by the work-item pass
    int32_t shared;  // another job of the launch writes the same grid: every
                     // tile is flushed with atomics
    // ordered reduction (gather / window, grids not shared): a split tile's
    // items write partial tiles, summed per pixel in item order afterwards,
    // so the grid is the same bit for bit on every run and no atomics are used
    int32_t det;
    int32_t taps;        // gridded by the small-support kernel (its own item queue)
    int32_t q2;          // a support-32 job beside other supports: the second owner queue
    int32_t pcap, rcap;  // partial-tile slots and reduce entries allocated
    void* partials;      // pcap tiles of T x T complex accumulators
    ReduceItem* reduce;  // rcap entries
};

// ---- host-side plan ---------------------------------------------------------
constexpr int kStageF64 = 0, kStageF32 = 1, kStageBf16 = 2;

// Everything that selects the owner-kernel instantiation; one launch serves
// jobs with equal keys.
struct KernelKey {
    int stage = kStageF64;  // staged coefficient type
    bool acc64 = false;     // float products summed in double
    int in_type = 0, out_type = 0;
    bool planar = false;
    int gm = 0;  // 0 table, 1 constant tap (COMPAT/coverage), 2 analytic
    int T = 32;
    bool operator==(const KernelKey& o) const
    {
        return stage == o.stage && acc64 == o.acc64 && in_type == o.in_type &&
               out_type == o.out_type && planar == o.planar && gm == o.gm && T == o.T;
    }
};

struct TileCfg {
    int T;   // tile width
    int Ty;  // tile height
    int H;
};
TileCfg pick_tile(int S, int tile_opt, bool taps);

struct TiledJob {
    GridArgs a;
    BinGeom bg;
    TileGeom g;
    bool rec64 = false;
    int vt = 0;
    int64_t e_base = 0;  // first element in the concatenated visibilities
    int64_t k_base = 0;  // first key (its keys are [k_base, k_base + nkeys])
    size_t off_split = 0, off_tcnt = 0;  // in the arena
    bool convert = false;  // the GCF table is converted to the staged type
    bool conv_dup = false;  // ... by an earlier job of the same table (shared copy)
    size_t off_gtab = 0;
    bool taps = false;  // small support: gridded by k_taps
    bool q2 = false;    // support 32 beside other supports: second owner queue and launch
};

// Keys and records of one job (device table, blockIdx.y of k_keyrec and
// k_gather): its elements are [e_base, e_base + npts) of the concatenated
// arrays, its keys [k_base, k_base + nkeys]; its binned visibilities are the
// sorted positions [offs[0], offs[nkeys]).
struct PackJob {
    const void* uv;
    const void* vals;
    const void* gcf;
    const int32_t* offs;
    int64_t e_base;
    int64_t npts;
    int64_t k_base;
    int64_t nkeys;
    int N, S, H, nb;
    int pitch;
    int esize;  // bytes of one staged element: record offsets are in bytes
};

// GCF table conversion to the staged type (device table).
struct ConvJob {
    const void* src;
    void* dst;
    int64_t n;          // complex elements written (64 s_dst^2)
    int32_t s_src;      // the caller's table support
    int32_t s_dst;      // the staged table's (s_src, or the support it is embedded in)
};

// Device layout of one owner launch over n jobs.  The jobs' visibilities
// are sorted together (one key space, each job a range of it): the arena
// holds the permutation, the CSR offsets and the records of all jobs, the
// per-job split flags and converted GCF tables, the job tables, the item
// queue and its state.
struct TiledPlan {
    KernelKey key;
    std::vector<TiledJob> jobs;
    int64_t E = 0;         // visibilities of all jobs
    int64_t K = 0;         // keys of all jobs
    int64_t max_npts = 0;  // the largest job
    int vt = 0;            // record value transform (shared)
    bool fresh = false;    // fresh grids (cleared by the plan's own passes)
    bool shared = false;   // jobs share grids (atomic flushes; the caller clears)
    size_t off_perm = 0, off_offs = 0, off_rec = 0;
    size_t off_rec0 = 0, off_pjobs = 0, off_cjobs = 0;
    int nconv = 0;             // jobs whose GCF is converted
    int64_t max_gcf = 0;       // largest GCF table (complex elements)
    size_t off_desc = 0, off_items = 0, off_qs = 0, arena_bytes = 0;
    size_t bin_bytes = 0;  // bin-sort scratch
    int64_t item_cap = 0;
    int64_t target = 0;    // visibilities per work item (0: one item per tile)
    int64_t target_t = 0;  // the same for the tap kernel's jobs
    int smax = 0;
    size_t lds = 0;     // the owner kernel's dynamic LDS (its staged sub-kernel layout)
    int stage_min = 0;  // tuning: staging threshold (0: default)
    int wg_per_cu = 0;  // tuning: persistent workgroups per CU (0: occupancy)
    bool det = false;   // ordered reduction of split tiles (gather / window)
    int64_t partial_cap = 0, reduce_cap = 0;
    size_t off_partials = 0, off_reduce = 0;
    // small supports go to the one-tap-per-lane kernel (k_taps), with their
    // own item queue; the other jobs to the owner kernel
    bool taps = false, owner = true;  // some job uses k_taps / the owner kernel
    int smax_owner = 0;               // the largest support among the owner kernel's jobs
    int s_common = 0;                 // their common support, 0 if they differ (-1: all 32 or 64)
    int s_common_t = 0;               // the same for the tap kernel's jobs
    size_t lds_taps = 0;
    size_t off_items_t = 0, off_qs_t = 0;
    int64_t item_cap_t = 0;
    // support-32 column-kernel jobs beside jobs of other supports: their own
    // queue, gridded by a second owner launch (the S = 32 instantiation)
    bool owner2 = false;
    size_t off_items_2 = 0, off_qs_2 = 0;
    int64_t item_cap_2 = 0;
};

struct OwnerLaunch {
    const JobDesc* jobs;
    const Item* items;
    QueueState* qs;
    int smax;
    int stage_min;  // staging threshold (0: default)
    int wg_per_cu;  // persistent workgroups per CU (0: occupancy limit)
    size_t lds;
    int sc = 0;  // every job's support when they share one (owner: 32, 64, or -1 for
                 // 32 and 64 chosen per item; taps: 8, 16)
};

int tiled_plan(const GridArgs* a, int n, const migrid_opts& o, TiledPlan& p);
void tiled_descs(const TiledPlan& p, char* arena, std::vector<JobDesc>& d);
// Writes the job table into the arena with kernels (graph-safe).
int tiled_write_descs(const TiledPlan& p, char* arena, const std::vector<JobDesc>& d,
                      hipStream_t s);
// Queue reset, GCF conversion, one bin sort of all jobs, record packing,
// per-tile counts and work items (one launch each).  For fresh grids the
// work-item pass zeroes the tiles the owner kernel will not store whole.
//
// Sorts can be kept between launches (the serving loop's jobs whose
// visibility coordinates repeat): jobs [n_sorted, n) of the plan are not
// sorted but take kept[j - n_sorted], a sort saved by tiled_keep_sort from an
// earlier launch of the same coordinates (n_sorted < 0: sort every job).
int tiled_prepare(const TiledPlan& p, char* arena, void* bin_scratch, hipStream_t s,
                  int n_sorted = -1, const void* const* kept = nullptr);
// Device bytes of a kept sort of job t (its permutation and CSR offsets).
size_t tiled_sort_bytes(const TiledJob& t);
// Saves job j's sort out of the arena into dst (tiled_sort_bytes bytes),
// numbered within the job, after tiled_prepare on the same stream.
int tiled_keep_sort(const TiledPlan& p, const char* arena, int j, void* dst, hipStream_t s);
// The persistent kernels over their queued items: the tap kernel for the
// small-support jobs, the owner kernel for the rest, then (gather / window)
// the ordered sum of split tiles.
int tiled_execute(const TiledPlan& p, char* arena, hipStream_t s);
int dispatch_owner(const KernelKey& k, const OwnerLaunch& L, hipStream_t s);
// The small-support kernel (one tap per lane, LDS tile) over all queued items.
int dispatch_taps(const KernelKey& k, const OwnerLaunch& L, hipStream_t s);

}  // namespace migrid
