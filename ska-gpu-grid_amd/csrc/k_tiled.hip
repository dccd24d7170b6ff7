// Tiled owner gridder: bin-sorted visibilities, sub-grid tiles held in
// registers, sub-kernels staged in LDS, atomics only where a tile is flushed.
//
// This is the build's form of all three reference strategies
// (scatter grid_gpu.cu:95-173, gather :391-503, moving window :504-673).
// Gridding is a weighted scatter, not a dense contraction, so MFMA is not
// used; the kernel is built for wave64 VALU + LDS:
//
//  * The grid is cut into T x T tiles (T = 16, 32, 64).  A workgroup of
//    (T/16)^2 waves owns one tile; each wave owns a 16 x 16 block of it, and
//    lane (lx, ly) = (lane & 7, lane >> 3) owns the 2 x 2 pixels
//    (Xw + 8j + lx, Yw + 8i + ly), accumulated in registers.
//  * Visibilities are bin-sorted by (k, bin_y, bin_x) over q = floor(uv)+S/2
//    in bins of 16 (k_bin.hip).  Pixel x receives visibilities with q_x in
//    [x, x+S-1], so a wave's visibilities for one k are a few contiguous CSR
//    ranges, and every one of them reaches the wave's block.
//  * For each k the workgroup stages sub-kernel k into LDS inside a zero
//    border wide enough for any listed visibility: each visibility costs a
//    lane four ds_reads at immediate offsets and 16 FMAs, with no per-lane
//    masks and no branches (taps outside the window read zeros).  A row pitch
//    of 8 mod 32 elements keeps the lane block's rows in distinct banks.  A k
//    with few visibilities for the workgroup skips staging and reads the
//    table from HBM with per-lane window masks.
//  * Visibility records (q, value) are wave-uniform and come through scalar
//    loads (constant address space); the complex multiply-accumulate is 4
//    FMAs per tap with the value in SGPRs.
//  * Work items are (tile, k-range) pairs sized by visibility count, taken
//    from a device queue by persistent workgroups.  A tile that is one item
//    is flushed with a plain read-add-write; a split tile with hardware
//    atomics (FP64, FP32, or FP32 with a two-sum compensation grid).
//  * Small supports (S <= 16, TILED strategy) run the tap kernel k_taps
//    instead, from their own item queue: lanes are taps of the window rather
//    than pixels of a block, and the tile is summed in LDS with atomics.
//
// Tap semantics as the reference (grid.cu:115-132): tap (r, c) of a
// visibility with q = main + S/2 lands on pixel (q_x - S + 1 + c,
// q_y - S + 1 + r) with coefficient G[k][r][c].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <utility>

#include "bin.hpp"
#include "device_common.hpp"
#include "internal.hpp"
#include "table.hpp"
#include "tiled.hpp"

namespace migrid {
namespace {


// ---- staged coefficient types ------------------------------------------------
struct Bf16x2 {
    uint32_t w;
};

__device__ __forceinline__ uint32_t f32_to_bf16_rne(float f)
{
    uint32_t b = __float_as_uint(f);
    b += 0x7fffu + ((b >> 16) & 1u);
    return b >> 16;
}

template <typename G>
struct Stage;
template <>
struct Stage<double2> {
    using P = double;
    __device__ static double2 make(double r, double i) { return make_double2(r, i); }
    __device__ static void get(const double2& g, double& r, double& i)
    {
        r = g.x;
        i = g.y;
    }
};
template <>
struct Stage<float2> {
    using P = float;
    __device__ static float2 make(double r, double i) { return make_float2((float)r, (float)i); }
    __device__ static void get(const float2& g, float& r, float& i)
    {
        r = g.x;
        i = g.y;
    }
};
template <>
struct Stage<Bf16x2> {
    using P = float;
    __device__ static Bf16x2 make(double r, double i)
    {
        Bf16x2 b;
        b.w = f32_to_bf16_rne((float)r) | (f32_to_bf16_rne((float)i) << 16);
        return b;
    }
    __device__ static void get(const Bf16x2& g, float& r, float& i)
    {
        r = __uint_as_float(g.w << 16);
        i = __uint_as_float(g.w & 0xffff0000u);
    }
};

template <typename T>
struct Cplx {
    T re, im;
};

// ---- packing ------------------------------------------------------------------
// Record value transforms.
constexpr int kValPlain = 0;    // the visibility value
constexpr int kValCompatS = 1;  // value x COMPUTE_GCF tap value, scatter form
constexpr int kValCompatI = 2;  // the same with the fixed-point dx of gather/window
constexpr int kValDebug = 3;    // (1, n): the DEBUG1 coverage map

// Keys and records of every job's visibilities in input order (blockIdx.y =
// job): the bin key (bin_key, the bin sort's key) into the sort's key buffer
// and the record -- q, its offset in the staged layout and the value, with
// the value transform of VT -- into rec0.  Both are read and written in
// order; the sorted records are then one gather (k_gather), one record read
// per visibility instead of its coordinates and value from two arrays.
template <typename TI, typename R, int VT>
__global__ __launch_bounds__(256) void k_keyrec(const PackJob* __restrict__ jobs,
                                                uint32_t* __restrict__ keys, R* __restrict__ rec0)
{
    const PackJob jb = jobs[blockIdx.y];
    const TI* __restrict__ uv = (const TI*)jb.uv;
    const TI* __restrict__ vals = (const TI*)jb.vals;
    const TI* __restrict__ gcf = (const TI*)jb.gcf;
    const int S = jb.S, pitch = jb.pitch, esize = jb.esize;
    const uint32_t nkeys = (uint32_t)jb.nkeys, kbase = (uint32_t)jb.k_base;
    float p1 = 0.f, p2 = 0.f;
    if (VT == kValCompatS || VT == kValCompatI) {
        const int64_t e = (int64_t)S * (S - 1) / 2 - 1;  // grid_gpu.cu:104-107
        const double T = (double)gcf[2 * e], w = (double)gcf[2 * e + 1];
        p1 = (float)(2 * 3.1415926 * w);
        p2 = (float)(p1 * T);
    }
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < jb.npts;
         n += (int64_t)gridDim.x * blockDim.x) {
        const TI u = uv[2 * n], v = uv[2 * n + 1];
        keys[jb.e_base + n] = kbase + bin_key<TI>(u, v, jb.N, S, jb.H, jb.nb, nkeys);
        // a visibility outside the grid keeps the sentinel key and its
        // record is never gathered
        const TI fu = floor_t(u), fv = floor_t(v);
        const int mx = (int)fu, my = (int)fv;
        const uint32_t qx = (uint32_t)(mx + S / 2), qy = (uint32_t)(my + S / 2);
        TI vr = vals[2 * n], vi = vals[2 * n + 1];
        if (VT == kValCompatS || VT == kValCompatI) {
            const int sx = (int)floor_t((TI)kOversample * (u - fu));
            const int sy = (int)floor_t((TI)kOversample * (v - fv));
            float xs, ys;
            if (VT == kValCompatS) {  // grid_gpu.cu:141-142, float uv
                xs = (float)(mx - u + sx * 1.0 / 8.0);
                ys = (float)(my - v + sy * 1.0 / 8.0);
            } else {  // grid_gpu.cu:456-459, int2 fixed point inn = main*8+sub
                const int ix = mx * kOversample + sx, iy = my * kOversample + sy;
                xs = (float)(mx - ix + sx * 1.0 / kOversample);
                ys = (float)(my - iy + sy * 1.0 / kOversample);
            }
            xs *= xs;
            ys *= ys;
            const float ph = p1 - p2 * sqrtf(xs + ys);
            float s_, c_;
            sincosf(ph, &s_, &c_);
            const TI r2 = (TI)s_, i2 = (TI)c_;
            const TI nr = vr * r2 - vi * i2;
            const TI ni = vr * i2 + r2 * vi;
            vr = nr;
            vi = ni;
        } else if (VT == kValDebug) {
            vr = (TI)1;
            vi = (TI)n;  // the visibility's index within its job
        }
        R r;
        r.q = qx | (qy << 16);
        r.off = (qy * (uint32_t)pitch + qx) * (uint32_t)esize;
        r.vr = (decltype(r.vr))vr;
        r.vi = (decltype(r.vi))vi;
        rec0[jb.e_base + n] = r;
    }
}

// Sorted records of every job (blockIdx.y = job): the job's binned
// visibilities are the sorted positions [offs[0], offs[nkeys]), and position
// j takes the record of element perm[j].
template <typename R>
__global__ __launch_bounds__(256) void k_gather(const PackJob* __restrict__ jobs,
                                                const int32_t* __restrict__ perm,
                                                const R* __restrict__ rec0, R* __restrict__ rec)
{
    const PackJob jb = jobs[blockIdx.y];
    const int64_t p_lo = jb.offs[0], p_hi = jb.offs[jb.nkeys];
    for (int64_t j = p_lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < p_hi;
         j += (int64_t)gridDim.x * blockDim.x)
        rec[j] = rec0[perm[j]];
}

// ---- per-tile counts ----------------------------------------------------------------
// tcount[tile * 64 + k] = visibilities of sub-kernel k whose window reaches the
// tile: the tile's bin rows summed from the CSR offsets.  Block (ty, k), one
// thread per tile column, so a bin row's offsets are read by neighbouring
// threads; blockIdx.y = job.
__global__ __launch_bounds__(128) void k_tile_counts(const JobDesc* __restrict__ jobs)
{
    const JobDesc& d = jobs[blockIdx.y];
    const TileGeom g = d.g;
    const int32_t* __restrict__ offs = d.offs;
    int32_t* __restrict__ tcount = (int32_t*)d.tcount;
    for (int64_t b = blockIdx.x; b < (int64_t)g.nty * kSubKernels; b += gridDim.x) {
        const int ty = (int)(b / kSubKernels), k = (int)(b % kSubKernels);
        const int byl = ty * g.Ty / g.H;
        const int byh = min(g.nb - 1, (ty * g.Ty + g.Ty + g.S - 2) / g.H);
        for (int tx = threadIdx.x; tx < g.ntx; tx += blockDim.x) {
            const int bxl = tx * g.T / g.H;
            const int bxh = min(g.nb - 1, (tx * g.T + g.T + g.S - 2) / g.H);
            int cnt = 0;
            for (int by = byl; by <= byh; by++) {
                const int64_t key = ((int64_t)k * g.nb + by) * g.nb;
                cnt += offs[key + bxh + 1] - offs[key + bxl];
            }
            tcount[((int64_t)ty * g.ntx + tx) * kSubKernels + k] = cnt;
        }
    }
}

// A pixel's (re, im) pair of an interleaved grid, stored or added as one
// 16- (double) or 8-byte (float) access: a wave's 8-pixel row segment then
// writes whole 128-byte lines, not half of each line per instruction.
template <typename TO>
struct Pair2;
template <>
struct Pair2<double> {
    using T = double2;
};
template <>
struct Pair2<float> {
    using T = float2;
};

// Zeroes the pixels of one tile (the whole wave; lanes along a row).
template <typename TO>
__device__ void zero_tile(TO* grid, int N, int X0, int Y0, int T, int Ty, bool planar)
{
    const int lane = threadIdx.x & 63;
    const int64_t NN = (int64_t)N * N;
    const int xe = min(N, X0 + T), ye = min(N, Y0 + Ty);
    for (int y = Y0; y < ye; y++)
        for (int x = X0 + lane; x < xe; x += 64) {
            const int64_t p = (int64_t)y * N + x;
            if (planar) {
                grid[p] = (TO)0;
                grid[NN + p] = (TO)0;
            } else {
                typename Pair2<TO>::T z;
                z.x = z.y = (TO)0;
                *(typename Pair2<TO>::T*)(grid + 2 * p) = z;
            }
        }
}

// ---- work items -----------------------------------------------------------------
// One wave per tile; lane = sub-kernel k.  Cuts the tile's sub-kernels into
// items of about `target` visibilities (target <= 0: one item per tile).
// Every job in one launch: blockIdx.y = job; a small-support job's items go
// to the small-support kernel's queue.  For a fresh grid the wave
// zeroes its tile when the owner kernel will not store it whole: no items
// (nothing stores it) or several (they add to it atomically).
template <typename TO>
__global__ __launch_bounds__(256) void k_items(const JobDesc* __restrict__ jobs, int64_t target_o,
                                               int64_t target_t, int planar,
                                               Item* __restrict__ items_o,
                                               QueueState* __restrict__ qs_o,
                                               Item* __restrict__ items_t,
                                               QueueState* __restrict__ qs_t,
                                               Item* __restrict__ items_2,
                                               QueueState* __restrict__ qs_2)
{
    const int job = blockIdx.y;
    // the owner kernel's queue, the small-support kernel's, or the second
    // owner queue; split tiles of the ordered reduction are listed in the
    // owner queue's counters whichever queue holds their items
    Item* __restrict__ items = jobs[job].taps ? items_t : jobs[job].q2 ? items_2 : items_o;
    QueueState* __restrict__ qs = jobs[job].taps ? qs_t : jobs[job].q2 ? qs_2 : qs_o;
    const int64_t target = jobs[job].taps ? target_t : target_o;
    const TileGeom g = jobs[job].g;
    const int32_t* __restrict__ tcount = jobs[job].tcount;
    uint8_t* __restrict__ split = (uint8_t*)jobs[job].split;
    const bool fresh = jobs[job].fresh != 0;
    const bool shared = jobs[job].shared != 0;
    const bool det = jobs[job].det != 0;
    const int lane = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t ntiles = (int64_t)g.ntx * g.nty;
    if (tile >= ntiles) return;
    const int tx = (int)(tile % g.ntx), ty = (int)(tile / g.ntx);
    const int64_t cnt = tcount[tile * kSubKernels + lane];
    // inclusive scan over the 64 lanes
    int64_t inc = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    const int64_t total = __shfl(inc, 63, 64);
    if (total == 0) {
        if (fresh) {
            zero_tile((TO*)jobs[job].grid, g.N, tx * g.T, ty * g.Ty, g.T, g.Ty, planar != 0);
            if (jobs[job].grid_lo)
                zero_tile((TO*)jobs[job].grid_lo, g.N, tx * g.T, ty * g.Ty, g.T, g.Ty, planar != 0);
        }
        return;
    }
    int nit = 1;
    if (target > 0) {
        int64_t want = (total + target - 1) / target;
        nit = (int)(want < 64 ? want : 64);
    }
    // item index of each non-empty k; an item starts at a non-empty k whose
    // index differs from that of the previous non-empty k, so no item is empty
    const int64_t excl = inc - cnt;
    const int j = (int)(excl * nit / total);
    const uint64_t nz = __ballot(cnt > 0);
    const uint64_t below = nz & ((1ull << lane) - 1);
    const int prev = below ? 63 - __builtin_clzll(below) : -1;
    const int jprev = __shfl(j, prev < 0 ? 0 : prev, 64);
    const bool start = cnt > 0 && (prev < 0 || j != jprev);
    const uint64_t m = __ballot(start);
    const int n_items = __popcll(m);
    int base = 0, pbase = -1;
    if (lane == 0) {
        base = atomicAdd(&qs->nitems, n_items);
        if (det && n_items > 1) {  // list the split tile for the ordered reduction
            const int rb = atomicAdd(&qs_o->nreduce, 1);
            const int pb = atomicAdd(&qs_o->npartial, n_items);
            // the plan's capacities bound every tile's need (tiled_plan), so
            // this always fits; were it not to, the tile would flush with
            // atomics (correct, not bitwise reproducible), counted in
            // qs->overflow
            const bool fits = rb < jobs[job].rcap && pb + n_items <= jobs[job].pcap;
            if (!fits) atomicAdd(&qs_o->overflow, 1);
            if (rb < jobs[job].rcap) {
                ReduceItem ri;
                ri.job = job;
                ri.tile = (int32_t)tile;
                ri.pslot = pb;
                ri.n = fits ? n_items : 0;
                jobs[job].reduce[rb] = ri;
            }
            if (fits) pbase = pb;
        }
    }
    base = __shfl(base, 0, 64);
    pbase = __shfl(pbase, 0, 64);
    if (start) {
        const int slot = base + __popcll(m & ((1ull << lane) - 1));
        const uint64_t above = lane == 63 ? 0 : (m >> (lane + 1));
        const int khi = above ? lane + __builtin_ctzll(above) : 63;
        Item it;
        it.tile = (int32_t)tile;
        it.klo = (int16_t)lane;
        it.khi = (int16_t)khi;
        it.job = job;
        it.pslot = pbase >= 0 ? pbase + __popcll(m & ((1ull << lane) - 1)) : -1;
        items[slot] = it;
    }
    if (lane == 0) split[tile] = n_items > 1 || shared;
    if (fresh && n_items > 1 && pbase < 0) {  // split tiles flushed with atomics start at zero
        zero_tile((TO*)jobs[job].grid, g.N, tx * g.T, ty * g.Ty, g.T, g.Ty, planar != 0);
        if (jobs[job].grid_lo)
            zero_tile((TO*)jobs[job].grid_lo, g.N, tx * g.T, ty * g.Ty, g.T, g.Ty, planar != 0);
    }
}

// ---- flush helpers --------------------------------------------------------------
// atomic: the tile is split over several items (adds with hardware atomics);
// otherwise the tile is this item's alone: stored outright when the grid is
// fresh, else read-add-written.
template <typename TO, typename TA>
__device__ __forceinline__ void flush_px(TO* grid, TO* lo, int64_t idx_re, int64_t idx_im, TA re,
                                         TA im, bool atomic, bool fresh)
{
    if (!atomic && lo == nullptr && idx_im == idx_re + 1) {
        // interleaved, the tile this item's alone
        using V2 = typename Pair2<TO>::T;
        V2* q = (V2*)(grid + idx_re);
        if (fresh) {
            // a fresh tile's pixels are not read again by this launch: a
            // nontemporal store, so they do not evict the staged tables
            // from L2 (6-9% fewer bytes fetched per launch)
            typedef TO E2 __attribute__((ext_vector_type(2)));
            E2 e;
            e.x = (TO)re;
            e.y = (TO)im;
            __builtin_nontemporal_store(e, (E2*)q);
            return;
        }
        V2 v = *q;
        v.x += (TO)re;
        v.y += (TO)im;
        *q = v;
        return;
    }
    if (!atomic && fresh) {
        grid[idx_re] = (TO)re;
        grid[idx_im] = (TO)im;
        if (lo != nullptr) {  // two-sum onto zero leaves no error
            lo[idx_re] = (TO)0;
            lo[idx_im] = (TO)0;
        }
        return;
    }
    if (lo != nullptr) {
        // FP32 grid with a two-sum compensation grid (Knuth's TwoSum on the
        // value the hardware added to): grid + lo carries the rounding error
        const float t[2] = {(float)re, (float)im};
        const int64_t ix[2] = {idx_re, idx_im};
        for (int c = 0; c < 2; c++) {
            float* hp = (float*)grid + ix[c];
            float a = atomic ? atomicAdd(hp, t[c]) : *hp;
            const float s = a + t[c];
            if (!atomic) *hp = s;
            const float bb = s - a;
            const float err = (a - (s - bb)) + (t[c] - bb);
            float* lp = (float*)lo + ix[c];
            if (atomic)
                atomic_add(lp, err);
            else
                *lp += err;
        }
        return;
    }
    if (atomic) {
        atomic_add(grid + idx_re, (TO)re);
        atomic_add(grid + idx_im, (TO)im);
    } else {
        grid[idx_re] += (TO)re;
        grid[idx_im] += (TO)im;
    }
}

// ---- the owner kernel -----------------------------------------------------------
// Zero border of a staged sub-kernel.  A wave covers a 16 x 16 block at a
// multiple of 16 and its visibility bins are 16 pixels wide, so a listed
// visibility's q lies 0 .. 16*floor((S+14)/16) + 15 pixels past the block's
// origin.  A lane at block offset o (0..15) then reads tap S - 1 + o - (q -
// origin), which runs from S - 1 - 16*floor((S+14)/16) - 15 up to S + 14: the
// border before the kernel (stage_border) covers the low end, 15 more rows
// and columns after it the high end, and every such read lands in the staged
// kernel or in zeros.  At S <= 64 the layout is at most 95 x 104 elements
// (158 KB of double pairs).
__host__ __device__ constexpr int stage_border(int S) { return 16 * ((S + 15 + 15) / 16) - S; }
__host__ __device__ constexpr int stage_rows(int S) { return S + stage_border(S) + 15; }
// Row pitch = 8 (mod 32) elements: the 8 rows of an 8 x 8 lane block fall in
// different LDS banks for 4-, 8- and 16-byte elements.
__host__ __device__ constexpr int stage_pitch(int S)
{
    return stage_rows(S) + ((8 - stage_rows(S) % 32) + 32) % 32;
}
constexpr int kStageMinDefault = 4;  // fewer visibilities for a k: read the table from HBM
// LDS of a workgroup beside the kernel's static arrays (160 KB per CU)
constexpr size_t kMaxLds = 160 * 1024 - 4096;

template <typename R>
using ConstPtr = const __attribute__((address_space(4))) R*;  // scalar-loadable

// Two consecutive records by one scalar load (s_load_dwordx8 for 16-byte
// records, x16 for 32-byte ones): the byte offset and value of each.
typedef uint32_t U32x8 __attribute__((ext_vector_type(8)));
typedef uint32_t U32x16 __attribute__((ext_vector_type(16)));
template <typename R>
struct Pair;
template <>
struct Pair<RecF32> {
    static __device__ __forceinline__ void get(ConstPtr<RecF32> rec, int r, uint32_t (&off)[2],
                                               float (&vr)[2], float (&vi)[2])
    {
        const U32x8 w = *(const __attribute__((address_space(4))) U32x8*)(rec + r);
#pragma unroll
        for (int j = 0; j < 2; j++) {
            off[j] = w[4 * j + 1];
            vr[j] = __uint_as_float(w[4 * j + 2]);
            vi[j] = __uint_as_float(w[4 * j + 3]);
        }
    }
};
template <>
struct Pair<RecF64> {
    static __device__ __forceinline__ void get(ConstPtr<RecF64> rec, int r, uint32_t (&off)[2],
                                               double (&vr)[2], double (&vi)[2])
    {
        const U32x16 w = *(const __attribute__((address_space(4))) U32x16*)(rec + r);
#pragma unroll
        for (int j = 0; j < 2; j++) {
            off[j] = w[8 * j + 1];
            vr[j] = __hiloint2double((int)w[8 * j + 5], (int)w[8 * j + 4]);
            vi[j] = __hiloint2double((int)w[8 * j + 7], (int)w[8 * j + 6]);
        }
    }
};
// The same from a byte offset into the records (32-bit: the scalar load
// takes it as its offset register), and one record alone.
template <typename R>
__device__ __forceinline__ void pair_at(ConstPtr<char> base, uint32_t o, uint32_t (&off)[2],
                                        decltype(+R::vr) (&vr)[2], decltype(+R::vr) (&vi)[2])
{
    Pair<R>::get((ConstPtr<R>)(base + o), 0, off, vr, vi);
}
template <typename R>
__device__ __forceinline__ void rec_at(ConstPtr<char> base, uint32_t o, uint32_t& off,
                                       decltype(+R::vr)& vr, decltype(+R::vr)& vi)
{
    ConstPtr<R> r = (ConstPtr<R>)(base + o);
    off = r->off;
    vr = r->vr;
    vi = r->vi;
}

template <typename R>
using GlobalPtr = const __attribute__((address_space(1))) R*;

// Element loads from the global address space (member-wise: the staged types
// cannot be copied from an address-space-qualified object).
template <typename G>
__device__ __forceinline__ G gload(GlobalPtr<G> p, int64_t i);
template <>
__device__ __forceinline__ Bf16x2 gload<Bf16x2>(GlobalPtr<Bf16x2> p, int64_t i)
{
    Bf16x2 b;
    b.w = p[i].w;
    return b;
}
template <>
__device__ __forceinline__ float2 gload<float2>(GlobalPtr<float2> p, int64_t i)
{
    return make_float2(p[i].x, p[i].y);
}
template <>
__device__ __forceinline__ double2 gload<double2>(GlobalPtr<double2> p, int64_t i)
{
    return make_double2(p[i].x, p[i].y);
}

// acc += v * g: 4 FMAs, or float products summed in double (<double2,float2>).
template <typename P, typename TA>
__device__ __forceinline__ void mac1(Cplx<TA>& acc, P vr, P vi, P gr, P gi)
{
    if (sizeof(P) == sizeof(TA)) {
        acc.re = fma((TA)vr, (TA)gr, acc.re);
        acc.re = fma(-(TA)vi, (TA)gi, acc.re);
        acc.im = fma((TA)vr, (TA)gi, acc.im);
        acc.im = fma((TA)vi, (TA)gr, acc.im);
    } else {
        acc.re += (TA)(vr * gr - vi * gi);
        acc.im += (TA)(vr * gi + gr * vi);
    }
}

// acc[i][j] += v * g[i][j] for the 2 x 2 pixels of a lane: 4 FMAs per tap, or
// float products summed in double for the <double2,float2> pair.
template <typename P, typename TA>
__device__ __forceinline__ void mac4(Cplx<TA> (&acc)[2][2], P vr, P vi, const P (&gr)[2][2],
                                     const P (&gi)[2][2])
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            if (sizeof(P) == sizeof(TA)) {
                acc[i][j].re = fma((TA)vr, (TA)gr[i][j], acc[i][j].re);
                acc[i][j].re = fma(-(TA)vi, (TA)gi[i][j], acc[i][j].re);
                acc[i][j].im = fma((TA)vr, (TA)gi[i][j], acc[i][j].im);
                acc[i][j].im = fma((TA)vi, (TA)gr[i][j], acc[i][j].im);
            } else {
                acc[i][j].re += (TA)(vr * gr[i][j] - vi * gi[i][j]);
                acc[i][j].im += (TA)(vr * gi[i][j] + gr[i][j] * vi);
            }
        }
}

// WX x WY waves own a (16 WX) x (16 WY) tile; lane (lx, ly) of wave (wx, wy)
// owns pixels (Xw + lx + 8j, Yw + ly + 8i), i, j in {0, 1}.  SC: every job's
// support as a compile-time constant (0: read per job), which makes the staged
// layout's border and pitch, and so the LDS offsets, constants.
template <int WX, int WY, typename G, typename TA, typename R, typename TI, typename TO,
          bool PLANAR, int GM, int SC>
// Register budget: two workgroups per CU need 8 waves per SIMD at T = 64
// (the 4- and 8-byte staged kernels leave room for two in LDS).
__global__ __launch_bounds__(64 * WX * WY)
__attribute__((amdgpu_waves_per_eu(sizeof(G) == 16 ? 4 : 8))) void k_owner(
    const JobDesc* __restrict__ jobs,
                                                        const Item* __restrict__ items,
                                                        QueueState* __restrict__ qs, int smax,
                                                        int stage_min)
{
    using P = typename Stage<G>::P;  // product type
    constexpr int NT = 64 * WX * WY;
    static_assert(((WX * WY) & (WX * WY - 1)) == 0, "a power-of-two number of waves");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    G* sg = (G*)smem_raw;
    __shared__ int s_item;
    __shared__ Cplx<double> s_fac[2][64];  // analytic separable factors

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // wave-uniform; the mask (an identity for tid < NT) tells the compiler
    // w < WX * WY, so the staging loops' row bounds fold away
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6) & (WX * WY - 1);
    const int wx = w % WX, wy = w / WX;
    const int lx = lane & 7, ly = lane >> 3;

    for (int e = tid; e < stage_rows(smax) * stage_pitch(smax); e += NT)
        sg[e] = Stage<G>::make(0.0, 0.0);
    int staged_s = -1;  // support of the layout now in LDS

    for (;;) {
        if (tid == 0) s_item = atomicAdd(&qs->head, 1);
        __syncthreads();
        const int it_idx = s_item;
        if (it_idx >= qs->nitems) break;
        const Item it = items[it_idx];
        const int tile = __builtin_amdgcn_readfirstlane(it.tile);
        const int klo = __builtin_amdgcn_readfirstlane(it.klo);
        const int khi = __builtin_amdgcn_readfirstlane(it.khi);
        const int job = __builtin_amdgcn_readfirstlane(it.job);
        const int pslot = __builtin_amdgcn_readfirstlane(it.pslot);
        const TileGeom g = jobs[job].g;
        ConstPtr<R> rec = (ConstPtr<R>)(uintptr_t)jobs[job].rec;
        const G* __restrict__ gtab = (const G*)jobs[job].gcf;  // table in the staged type
        TO* __restrict__ grid = (TO*)jobs[job].grid;
        TO* __restrict__ grid_lo = (TO*)jobs[job].grid_lo;
        const double sigma = jobs[job].sigma, beta = jobs[job].beta;
        const bool atomic = jobs[job].split[tile] != 0;
        const bool fresh = jobs[job].fresh != 0;
        const int S = SC ? SC : g.S;
        const int B = stage_border(S);
        const int pitch = stage_pitch(S);
        const int64_t NN = (int64_t)g.N * g.N;
        if (S != staged_s) {
            // a new layout: clear the old interior, then (constant-tap form)
            // write the all-ones kernel once
            __syncthreads();
            if (staged_s > 0) {
                const int ob = stage_border(staged_s), op = stage_pitch(staged_s);
                for (int e = tid; e < staged_s * staged_s; e += NT)
                    sg[(e / staged_s + ob) * op + e % staged_s + ob] = Stage<G>::make(0.0, 0.0);
                __syncthreads();
            }
            if (GM == 1)
                for (int e = tid; e < S * S; e += NT)
                    sg[(e / S + B) * pitch + e % S + B] = Stage<G>::make(1.0, 0.0);
            staged_s = S;
        }
        __syncthreads();
        const int tx = tile % g.ntx, ty = tile / g.ntx;
        const int X0 = tx * g.T, Y0 = ty * g.T;
        const int Xw = X0 + 16 * wx, Yw = Y0 + 16 * wy;
        // this wave's bins (q in [Xw, Xw + S + 14]) and the workgroup's
        const int bxl = Xw / g.H, bxh = min(g.nb - 1, (Xw + S + 14) / g.H);
        const int byl = Yw / g.H, byh = min(g.nb - 1, (Yw + S + 14) / g.H);
        const int lane_off = ly * pitch + lx;

        Cplx<TA> acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j].re = acc[i][j].im = (TA)0;

        // lane k counts the workgroup's visibilities for sub-kernel k; the
        // ballots give the non-empty and the staged sub-kernels of this item
        uint64_t nz_mask, st_mask;
        {
            const int cnt =
                (lane >= klo && lane <= khi) ? jobs[job].tcount[(int64_t)tile * kSubKernels + lane] : 0;
            nz_mask = __ballot(cnt > 0);
            st_mask = GM != 0 ? nz_mask : __ballot(cnt >= stage_min && cnt > 0);
        }
        // sub-kernel rows this wave stages: w, w + NW, ... (one column per lane);
        // the next staged sub-kernel is fetched into registers while the
        // current one is in use
        constexpr int NW = WX * WY;
        constexpr bool PF = GM == 0 && NW >= 8;
        constexpr int PR = PF ? (64 + NW - 1) / NW : 1;
        G pf[PR];
        auto fetch_rows = [&](int kk) {
            GlobalPtr<G> src = (GlobalPtr<G>)(uintptr_t)(gtab + (int64_t)kk * S * S);
#pragma unroll
            for (int p = 0; p < PR; p++) {
                const int r = w + p * NW;
                if (r < S && lane < S) pf[p] = gload<G>(src, r * S + lane);
            }
        };
        if (PF && st_mask) fetch_rows(__builtin_ctzll(st_mask));

        // CSR bounds of this wave's bin rows for one sub-kernel, gathered by
        // one vector load (lane 2t: start of row byl + t, lane 2t+1: its end)
        // and fetched one sub-kernel ahead, so no scalar-memory round trip
        // sits between consecutive sub-kernels
        constexpr int NBY = 5;  // bin rows per wave for S <= 64
        GlobalPtr<int32_t> offs_v = (GlobalPtr<int32_t>)(uintptr_t)jobs[job].offs;
        // A wave whose block lies past the grid edge (an edge tile) has no
        // bins (its first bin index would be past the last one): its ranges
        // stay empty, and it still takes part in the workgroup's barriers.
        const bool wave_on = Xw < g.N && Yw < g.N && bxl <= bxh && byl <= byh;
        auto load_bounds = [&](int kk) -> int {
            int v = 0;
            const int t = lane >> 1;
            if (wave_on && lane < 2 * NBY && byl + t <= byh) {
                // 32-bit: keys stay below 2^31 (the plan refuses more)
                const int key = (kk * g.nb + byl + t) * g.nb;
                v = offs_v[(lane & 1) ? key + bxh + 1 : key + bxl];
            }
            return v;
        };
        int bnd_next = nz_mask ? load_bounds(__builtin_ctzll(nz_mask)) : 0;

        for (uint64_t todo = nz_mask; todo; todo &= todo - 1) {
            const int k = __builtin_ctzll(todo);
            const bool staged = (st_mask >> k) & 1;
            const int bnd = bnd_next;
            {
                const uint64_t rest = todo & (todo - 1);
                if (rest) bnd_next = load_bounds(__builtin_ctzll(rest));
            }
            if (staged && GM != 1) {
                __syncthreads();  // previous sub-kernel no longer read
                if (PF) {
#pragma unroll
                    for (int p = 0; p < PR; p++) {
                        const int r = w + p * NW;
                        if (r < S && lane < S) sg[(r + B) * pitch + lane + B] = pf[p];
                    }
                    const uint64_t later = st_mask & ~((2ull << k) - 1);
                    if (later) fetch_rows(__builtin_ctzll(later));
                } else if (GM == 0) {
                    // wave w stages rows w, w + NW, ...; lanes walk the columns
                    GlobalPtr<G> gk = (GlobalPtr<G>)(uintptr_t)(gtab + (int64_t)k * S * S);
                    for (int r = w; r < S; r += NW)
                        for (int c = lane; c < S; c += 64)
                            sg[(r + B) * pitch + c + B] = gload<G>(gk, r * S + c);
                } else {
                    // separable analytic GCF: G[r][c] = f(dy_r) * f(dx_c),
                    // f(d) = exp(-d^2/(2 sigma^2)) * exp(i beta d^2)
                    const int sx = k % kOversample, sy = k / kOversample;
                    for (int e = tid; e < 2 * S; e += NT) {
                        const int axis = e / S, t = e % S;
                        const double d = (double)(t - S / 2 + 1) - (axis ? sy : sx) / 8.0;
                        const double q = d * d;
                        const double amp = exp(-q / (2.0 * sigma * sigma));
                        double sn, cs;
                        sincos(beta * q, &sn, &cs);
                        s_fac[axis][t].re = amp * cs;
                        s_fac[axis][t].im = amp * sn;
                    }
                    __syncthreads();
                    for (int e = tid; e < S * S; e += NT) {
                        const int r = e / S, c = e % S;
                        const Cplx<double> a = s_fac[1][r], b = s_fac[0][c];
                        sg[(r + B) * pitch + c + B] =
                            Stage<G>::make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
                    }
                }
                __syncthreads();
            }
            if (staged) {
                // Per lane: the staged element of block (0,0) for q = 0; a
                // visibility shifts it back by its wave-uniform record offset
                // (q_y * pitch + q_x elements, stored in bytes), and taps
                // outside the window read the zero border.  The wave's bin
                // rows in turn, in one loop: pairs of visibilities from one
                // scalar record load, then a row's odd last one alone.  (One
                // loop over the rows keeps the accumulators in the same
                // registers throughout; a loop per row costs copies in and out
                // of each.)
                const char* lb = (const char*)(sg + lane_off + (S - 1 + B + Yw) * pitch +
                                               (S - 1 + B + Xw));
                auto taps = [&](uint32_t off, P (&gr)[2][2], P (&gi)[2][2]) {
                    const G* b = (const G*)(lb - off);
#pragma unroll
                    for (int i = 0; i < 2; i++)
#pragma unroll
                        for (int j = 0; j < 2; j++)
                            Stage<G>::get(b[8 * i * pitch + 8 * j], gr[i][j], gi[i][j]);
                };
                int t = 0;
                int r = __builtin_amdgcn_readlane(bnd, 0), e = __builtin_amdgcn_readlane(bnd, 1);
                for (;;) {
                    for (; r + 1 < e; r += 2) {
                        uint32_t off[2];
                        decltype(+rec[0].vr) vr[2], vi[2];
                        Pair<R>::get(rec, r, off, vr, vi);
                        P gr0[2][2], gi0[2][2], gr1[2][2], gi1[2][2];
                        taps(off[0], gr0, gi0);
                        taps(off[1], gr1, gi1);
                        mac4<P, TA>(acc, (P)vr[0], (P)vi[0], gr0, gi0);
                        mac4<P, TA>(acc, (P)vr[1], (P)vi[1], gr1, gi1);
                    }
                    if (r < e) {
                        P gr0[2][2], gi0[2][2];
                        taps(rec[r].off, gr0, gi0);
                        mac4<P, TA>(acc, (P)rec[r].vr, (P)rec[r].vi, gr0, gi0);
                    }
                    if (++t == NBY) break;
                    r = __builtin_amdgcn_readlane(bnd, 2 * t);
                    e = __builtin_amdgcn_readlane(bnd, 2 * t + 1);
                }
            } else {
                // a sparse sub-kernel read from the table in memory
                GlobalPtr<G> gk = (GlobalPtr<G>)(uintptr_t)(gtab + (int64_t)k * S * S);
#pragma unroll
                for (int t = 0; t < NBY; t++) {
                    const int r0 = __builtin_amdgcn_readlane(bnd, 2 * t);
                    const int r1 = __builtin_amdgcn_readlane(bnd, 2 * t + 1);
                    for (int r = r0; r < r1; r++) {
                        const uint32_t q = rec[r].q;
                        const int c0 = S - 1 - ((int)(q & 0xffffu) - Xw);
                        const int rr0 = S - 1 - ((int)(q >> 16) - Yw);
                        P gr[2][2], gi[2][2];
#pragma unroll
                        for (int i = 0; i < 2; i++)
#pragma unroll
                            for (int j = 0; j < 2; j++) {
                                const int c = c0 + lx + 8 * j, rr = rr0 + ly + 8 * i;
                                gr[i][j] = gi[i][j] = (P)0;
                                if ((unsigned)c < (unsigned)S && (unsigned)rr < (unsigned)S) {
                                    const int64_t t = (int64_t)rr * S + c;
                                    Stage<G>::get(gload<G>(gk, t), gr[i][j], gi[i][j]);
                                }
                            }
                        mac4<P, TA>(acc, (P)rec[r].vr, (P)rec[r].vi, gr, gi);
                    }
                }
            }
        }
        // flush: a split tile of the ordered-reduction mode writes its
        // partial tile (T x T accumulators) to its slot
        if (pslot >= 0) {
            TA* part = (TA*)jobs[job].partials + (int64_t)pslot * g.T * g.T * 2;
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const int e = (Yw - Y0 + 8 * i + ly) * g.T + (Xw - X0 + 8 * j + lx);
                    part[2 * e] = acc[i][j].re;
                    part[2 * e + 1] = acc[i][j].im;
                }
            __syncthreads();  // s_item reuse
            continue;
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int y = Yw + 8 * i + ly;
            if (y >= g.N) continue;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int x = Xw + 8 * j + lx;
                if (x >= g.N) continue;
                const int64_t p = (int64_t)y * g.N + x;
                if (PLANAR)
                    flush_px(grid, grid_lo, p, NN + p, acc[i][j].re, acc[i][j].im, atomic, fresh);
                else
                    flush_px(grid, grid_lo, 2 * p, 2 * p + 1, acc[i][j].re, acc[i][j].im, atomic,
                             fresh);
            }
        }
        __syncthreads();  // s_item reuse
    }
}

// ---- the column-block owner kernel (128 x 64 tiles) ----------------------------
// A workgroup of 16 waves owns a 128 x 64 tile as 16 blocks of 8 x 64 pixels:
// wave w's block starts at (X0 + 8 w, Y0), and lane (lx, ly) = (lane & 7,
// lane >> 3) owns the eight pixels (Xw + lx, Yw + ly + 8 i), i = 0..7, of its
// column, one in each row group of 8.  Bins are 8 x 8 pixels of q, so a wave
// lists col_ncols(S) bins in each of its col_nrows(S) bin rows, and every
// visibility of bin row t reaches the same row groups [max(0, t - D),
// min(7, t)], D = col_depth(S): a lane reads and multiplies only those.  Per
// axis an S-wide window then costs about S + 7 executed columns and rows,
// against S + 15 for 16 x 16 blocks (at S = 64, 81% of the executed taps are
// real rather than 66%); a visibility's taps need one address per lane, the
// groups being immediate offsets apart; and a 64-row block meets a window in
// two blocks' worth of rows at most, so each visibility is visited by about
// 18 waves rather than 25.
//
// Staged layout: sub-kernel element (r, c) sits at LDS element
// (r + L) P + c + L, with L = col_lead(S) zero rows and columns before the
// kernel and 7 after it (a row's trailing zeros are the next row's leading
// ones, so P >= S + L).  A lane reads row r = ly + 8 i - dy + S - 1, with
// dy = q_y - Yw in [8t, 8t + 7] and i one of bin row t's groups, and column
// c = lx - dx + S - 1, with dx = q_x - Xw in [0, 8 col_ncols(S) - 1]: both lie
// in [-L, S + 6], so every read lands in the kernel or in zeros.
constexpr int kColGroups = 8;  // row groups per lane
constexpr int kColTx = 128, kColTy = 64;  // tile width and height
constexpr int kColH = 8;                  // bin edge of the column kernel's jobs
__host__ __device__ constexpr int col_depth(int S) { return (S + 6) / 8; }
__host__ __device__ constexpr int col_lead(int S) { return 8 * col_depth(S) + 8 - S; }
// S + L is a multiple of 8; the pitch is 8 or 24 mod 32 elements, so the 8
// rows of an 8 x 8 lane block fall in different LDS banks, and a multiple of
// 8, so row groups (8 P elements apart) pair up in ds_read2st64 reads.
__host__ __device__ constexpr int col_pitch(int S)
{
    return (S + col_lead(S)) % 16 == 8 ? S + col_lead(S) : S + col_lead(S) + 8;
}
__host__ __device__ constexpr int col_rows(int S) { return S + col_lead(S) + 8; }
__host__ __device__ constexpr int col_ncols(int S) { return col_depth(S) + 1; }
__host__ __device__ constexpr int col_nrows(int S) { return (8 * kColGroups + S - 2) / 8 + 1; }

template <int... I, typename F>
__device__ __forceinline__ void static_for_seq(std::integer_sequence<int, I...>, F&& f)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_seq(std::make_integer_sequence<int, N>{}, f);
}

// SC: every job's support as a compile-time constant (0: read per job), which
// makes the bin rows' group ranges and the LDS offsets constants; kScBoth:
// every job at 64 or 32, the constant chosen per item.
constexpr int kScBoth = -1;

template <typename G, typename TA, typename R, typename TI, typename TO, bool PLANAR, int GM,
          int SC>
__global__ __launch_bounds__(1024)
__attribute__((amdgpu_waves_per_eu(sizeof(G) == 16 ? 4 : 8))) void k_owner_col(
    const JobDesc* __restrict__ jobs, const Item* __restrict__ items,
    QueueState* __restrict__ qs, int smax)
{
    using P = typename Stage<G>::P;  // product type
    using V = decltype(+((const R*)nullptr)->vr);  // record value type
    constexpr int NW = 16, NT = 64 * NW, NG = kColGroups;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    G* sg = (G*)smem_raw;
    __shared__ int s_item;
    __shared__ Cplx<double> s_fac[2][64];  // analytic separable factors

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6) & (NW - 1);
    const int lx = lane & 7, ly = lane >> 3;

    // (two support-32 layouts side by side when S = 32 sub-kernels are staged in pairs)
    const int lds_elems = max(col_rows(smax) * col_pitch(smax),
                              smax >= 32 ? 2 * col_rows(32) * col_pitch(32) : 0);
    for (int e = tid; e < lds_elems; e += NT)
        sg[e] = Stage<G>::make(0.0, 0.0);
    int staged_s = -1;  // support of the layout now in LDS

    for (;;) {
        if (tid == 0) s_item = atomicAdd(&qs->head, 1);
        __syncthreads();
        const int it_idx = s_item;
        if (it_idx >= qs->nitems) break;
        const Item it = items[it_idx];
        const int tile = __builtin_amdgcn_readfirstlane(it.tile);
        const int klo = __builtin_amdgcn_readfirstlane(it.klo);
        const int khi = __builtin_amdgcn_readfirstlane(it.khi);
        const int job = __builtin_amdgcn_readfirstlane(it.job);
        const int pslot = __builtin_amdgcn_readfirstlane(it.pslot);
        const TileGeom g = jobs[job].g;
        ConstPtr<R> rec = (ConstPtr<R>)(uintptr_t)jobs[job].rec;
        ConstPtr<int32_t> offs = (ConstPtr<int32_t>)(uintptr_t)jobs[job].offs;
        const G* __restrict__ gtab = (const G*)jobs[job].gcf;  // table in the staged type
        TO* __restrict__ grid = (TO*)jobs[job].grid;
        TO* __restrict__ grid_lo = (TO*)jobs[job].grid_lo;
        const double sigma = jobs[job].sigma, beta = jobs[job].beta;
        const bool atomic = jobs[job].split[tile] != 0;
        const bool fresh = jobs[job].fresh != 0;
        // the item's work with its support SX a compile-time constant (0: read
        // from the job); SC = kScBoth picks 64 or 32 per item
        auto body = [&](auto scx) {
            constexpr int SX = decltype(scx)::value;
            const int S = SX ? SX : g.S;
            const int L = col_lead(S);
            const int pitch = col_pitch(S);
            const int64_t NN = (int64_t)g.N * g.N;
            if (S != staged_s) {
                // a new layout: clear the old interior, then (constant-tap form)
                // write the all-ones kernel once
                __syncthreads();
                if (staged_s > 0) {
                    const int ol = col_lead(staged_s), op = col_pitch(staged_s);
                    // (the second of a pair's layouts too, at S = 32)
                    const int nl = GM == 0 && staged_s == 32 ? 2 : 1;
                    for (int e = tid; e < nl * staged_s * staged_s; e += NT) {
                        const int l = e / (staged_s * staged_s), f = e % (staged_s * staged_s);
                        sg[l * col_rows(32) * col_pitch(32) + (f / staged_s + ol) * op + f % staged_s + ol] =
                            Stage<G>::make(0.0, 0.0);
                    }
                    __syncthreads();
                }
                if (GM == 1)
                    for (int e = tid; e < S * S; e += NT)
                        sg[(e / S + L) * pitch + e % S + L] = Stage<G>::make(1.0, 0.0);
                staged_s = S;
            }
            __syncthreads();
            const int tx = tile % g.ntx, ty = tile / g.ntx;
            const int X0 = tx * kColTx, Y0 = ty * kColTy;
            const int Xw = X0 + 8 * w, Yw = Y0;
            const int nb = g.nb;
            // this wave's bins: columns bxl .. bxl + ncx - 1, rows byl .. byl + nry - 1
            // (fewer at the grid's far edge; none for a block past it, whose wave
            // still takes part in the workgroup's barriers)
            const int bxl = Xw / kColH, byl = Yw / kColH;
            const bool on = Xw < g.N && Yw < g.N;
            const int ncx = on ? min(nb - bxl, col_ncols(S)) : 0;
            const int nry = on && ncx > 0 ? min(nb - byl, col_nrows(S)) : 0;
            const int D = col_depth(S);

            Cplx<TA> acc[NG];
    #pragma unroll
            for (int i = 0; i < NG; i++) acc[i].re = acc[i].im = (TA)0;

            // lane k counts the workgroup's visibilities for sub-kernel k; the
            // ballots give the non-empty and the staged sub-kernels of this item
            uint64_t nz_mask, st_mask;
            {
                const int cnt = (lane >= klo && lane <= khi)
                                    ? jobs[job].tcount[(int64_t)tile * kSubKernels + lane]
                                    : 0;
                nz_mask = __ballot(cnt > 0);
                // every non-empty sub-kernel is staged (table jobs always run
                // with the support a constant, whose registers a direct-from-
                // memory path would take)
                st_mask = nz_mask;
            }
            // sub-kernel rows this wave stages: w, w + NW, ... (one column per
            // lane); the next staged sub-kernel is fetched into registers while
            // the current one is in use
            constexpr bool PF = GM == 0;
            constexpr int PR = PF ? (64 + NW - 1) / NW : 1;
            G pf[PR];
            auto fetch_rows = [&](int kk) {
                GlobalPtr<G> src = (GlobalPtr<G>)(uintptr_t)(gtab + (int64_t)kk * S * S);
    #pragma unroll
                for (int p = 0; p < PR; p++) {
                    const int r = w + p * NW;
                    if (r < S && lane < S) pf[p] = gload<G>(src, r * S + lane);
                }
            };
            // S = 32 (table): two sub-kernels staged side by side per barrier
            // pair (layouts at 0 and col_rows(32) col_pitch(32) elements); the
            // four prefetch registers hold both, two rows of each per wave
            constexpr bool PAIR = SX == 32 && GM == 0;
            if (PF && !PAIR && st_mask) fetch_rows(__builtin_ctzll(st_mask));

            // Per lane: the staged element of group 0 for q = 0, in bytes; a
            // visibility shifts it back by its record offset (q_y P + q_x
            // elements, in bytes), group i lies 8 i P elements further.
            const char* lb = (const char*)sg +
                             (size_t)((Yw + ly + S - 1 + L) * pitch + Xw + lx + S - 1 + L) * sizeof(G);

            // CSR bounds of the wave's bin rows for one sub-kernel, gathered by
            // one vector load (lane 2t: start of row t, lane 2t + 1: its end) and
            // fetched one sub-kernel ahead; row t's key is key0 + t nb
            GlobalPtr<int32_t> offs_v = (GlobalPtr<int32_t>)(uintptr_t)jobs[job].offs;
            auto load_bounds = [&](int kk) -> int {
                int v = 0;
                if (lane < 2 * nry) {
                    const int key = (kk * nb + byl + (lane >> 1)) * nb + bxl;
                    v = offs_v[(lane & 1) ? key + ncx : key];
                }
                return v;
            };
            int bnd_next = nz_mask ? load_bounds(__builtin_ctzll(nz_mask)) : 0;

            // the rows of one sub-kernel, staged at lbx, bounds bndx, with the
            // support a constant SX: each bin row's groups and the group
            // offsets are constants
            auto rows_const = [&](auto sx, const char* lbx, int bndx) {
                constexpr int SX = decltype(sx)::value;
                constexpr int NR = col_nrows(SX), DC = col_depth(SX), PC = col_pitch(SX);
                constexpr uint32_t RS = sizeof(R);
                ConstPtr<char> rbase = (ConstPtr<char>)rec;
                // the bounds as byte offsets into the records, and a mask
                // of the non-empty rows (bit 2t: row t), so an empty row
                // costs one bit test
                const uint32_t bb = (uint32_t)bndx * RS;
                const uint32_t bb_next =
                    (uint32_t)__builtin_amdgcn_mov_dpp((int)bb, 0xB1, 0xF, 0xF, false);
                const uint64_t rows_on = __ballot(!(lane & 1) && bb < bb_next);
                static_for<NR>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int A = t - DC > 0 ? t - DC : 0;
                    constexpr int B = t < NG - 1 ? t : NG - 1;
                    if (!((rows_on >> (2 * t)) & 1)) return;
                    uint32_t o = (uint32_t)__builtin_amdgcn_readlane((int)bb, 2 * t);
                    const uint32_t oe = (uint32_t)__builtin_amdgcn_readlane((int)bb, 2 * t + 1);
                    // the taps of one visibility
                    auto one = [&](uint32_t off, V vr, V vi) {
                        const G* b = (const G*)(lbx - off);
                        P gr[NG], gi[NG];
#pragma unroll
                        for (int i = A; i <= B; i++) Stage<G>::get(b[8 * i * PC], gr[i], gi[i]);
#pragma unroll
                        for (int i = A; i <= B; i++)
                            mac1<P, TA>(acc[i], (P)vr, (P)vi, gr[i], gi[i]);
                    };
                    // pairs of visibilities from one scalar record load,
                    // then the row's odd last one alone
                    const uint32_t oe1 = oe - RS;
                    for (; o < oe1; o += 2 * RS) {
                        uint32_t off[2];
                        V vr[2], vi[2];
                        pair_at<R>(rbase, o, off, vr, vi);
                        one(off[0], vr[0], vi[0]);
                        // rows of many groups: the second visibility's
                        // taps are read after the first one's products,
                        // so one visibility's taps are live at a time
                        if constexpr (B - A + 1 > 4) __builtin_amdgcn_sched_barrier(0);
                        one(off[1], vr[1], vi[1]);
                    }
                    if (o == oe1) {
                        uint32_t off;
                        V vr, vi;
                        rec_at<R>(rbase, o, off, vr, vi);
                        one(off, vr, vi);
                    }
                });
            };

            if constexpr (PAIR) {
                constexpr int kPairOff = col_rows(32) * col_pitch(32);
                constexpr int PRS = PR / 2;
                auto fetch2 = [&](int kk, int slot) {
                    GlobalPtr<G> src = (GlobalPtr<G>)(uintptr_t)(gtab + (int64_t)kk * 32 * 32);
    #pragma unroll
                    for (int p = 0; p < PRS; p++) {
                        const int r = w + p * NW;
                        if (lane < 32) pf[slot * PRS + p] = gload<G>(src, r * 32 + lane);
                    }
                };
                const char* lb2 = lb + (size_t)kPairOff * sizeof(G);
                uint64_t todo = nz_mask;
                int b1 = 0, b2 = 0;
                if (todo) {
                    uint64_t t = todo;
                    const int k1 = __builtin_ctzll(t);
                    fetch2(k1, 0);
                    b1 = load_bounds(k1);
                    t &= t - 1;
                    if (t) {
                        fetch2(__builtin_ctzll(t), 1);
                        b2 = load_bounds(__builtin_ctzll(t));
                    }
                }
                while (todo) {
                    todo &= todo - 1;
                    const bool has2 = todo != 0;
                    if (has2) todo &= todo - 1;
                    const int bnd1 = b1, bnd2 = b2;
                    {
                        uint64_t t = todo;
                        if (t) {
                            b1 = load_bounds(__builtin_ctzll(t));
                            t &= t - 1;
                            if (t) b2 = load_bounds(__builtin_ctzll(t));
                        }
                    }
                    __syncthreads();  // the previous pair no longer read
    #pragma unroll
                    for (int p = 0; p < PRS; p++) {
                        const int r = w + p * NW;
                        if (lane < 32) {
                            sg[(r + L) * pitch + lane + L] = pf[p];
                            if (has2) sg[kPairOff + (r + L) * pitch + lane + L] = pf[PRS + p];
                        }
                    }
                    {
                        uint64_t t = todo;
                        if (t) {
                            fetch2(__builtin_ctzll(t), 0);
                            t &= t - 1;
                            if (t) fetch2(__builtin_ctzll(t), 1);
                        }
                    }
                    __syncthreads();  // both staged
                    rows_const(std::integral_constant<int, 32>{}, lb, bnd1);
                    if (has2) rows_const(std::integral_constant<int, 32>{}, lb2, bnd2);
                }
            } else {
            for (uint64_t todo = nz_mask; todo; todo &= todo - 1) {
                const int k = __builtin_ctzll(todo);
                const bool staged = (st_mask >> k) & 1;
                // first key of the wave's bin rows: row t at key0 + t nb
                const int key0 = (k * nb + byl) * nb + bxl;
                const int bnd = bnd_next;
                {
                    const uint64_t rest = todo & (todo - 1);
                    if (rest) bnd_next = load_bounds(__builtin_ctzll(rest));
                }
                if (staged && GM != 1) {
                    __syncthreads();  // previous sub-kernel no longer read
                    if (PF) {
    #pragma unroll
                        for (int p = 0; p < PR; p++) {
                            const int r = w + p * NW;
                            if (r < S && lane < S) sg[(r + L) * pitch + lane + L] = pf[p];
                        }
                        const uint64_t later = st_mask & ~((2ull << k) - 1);
                        if (later) fetch_rows(__builtin_ctzll(later));
                    } else {
                        // separable analytic GCF: G[r][c] = f(dy_r) * f(dx_c),
                        // f(d) = exp(-d^2/(2 sigma^2)) * exp(i beta d^2)
                        const int sx = k % kOversample, sy = k / kOversample;
                        for (int e = tid; e < 2 * S; e += NT) {
                            const int axis = e / S, t = e % S;
                            const double d = (double)(t - S / 2 + 1) - (axis ? sy : sx) / 8.0;
                            const double q = d * d;
                            const double amp = exp(-q / (2.0 * sigma * sigma));
                            double sn, cs;
                            sincos(beta * q, &sn, &cs);
                            s_fac[axis][t].re = amp * cs;
                            s_fac[axis][t].im = amp * sn;
                        }
                        __syncthreads();
                        for (int e = tid; e < S * S; e += NT) {
                            const int r = e / S, c = e % S;
                            const Cplx<double> a = s_fac[1][r], b = s_fac[0][c];
                            sg[(r + L) * pitch + c + L] =
                                Stage<G>::make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
                        }
                    }
                    __syncthreads();
                }
                if (staged) {
                    if constexpr (SX != 0) {
                        rows_const(std::integral_constant<int, SX>{}, lb, bnd);
                    } else if (S == 64) {
                        // jobs of mixed supports: the common ones take the
                        // constant-support rows, decided once per item
                        rows_const(std::integral_constant<int, 64>{}, lb, bnd);
                    } else if (S == 32) {
                        rows_const(std::integral_constant<int, 32>{}, lb, bnd);
                    } else {
                        // any support: the groups of bin row t are tested per
                        // visibility (wave-uniform branches)
                        for (int t = 0; t < nry; t++) {
                            const int A = max(0, t - D), B = min(NG - 1, t);
                            const int r0 = offs[key0 + t * nb], r1 = offs[key0 + t * nb + ncx];
                            for (int r = r0; r < r1; r++) {
                                const G* b = (const G*)(lb - rec[r].off);
                                const P vr = (P)rec[r].vr, vi = (P)rec[r].vi;
    #pragma unroll
                                for (int i = 0; i < NG; i++) {
                                    if (i < A || i > B) continue;
                                    P gr, gi;
                                    Stage<G>::get(b[8 * i * pitch], gr, gi);
                                    mac1<P, TA>(acc[i], vr, vi, gr, gi);
                                }
                            }
                        }
                    }
                }
            }
            }  // PAIR
            // flush: a split tile of the ordered-reduction mode writes its
            // partial tile (128 x 64 accumulators) to its slot
            if (pslot >= 0) {
                TA* part = (TA*)jobs[job].partials + (int64_t)pslot * kColTx * kColTy * 2;
    #pragma unroll
                for (int i = 0; i < NG; i++) {
                    const int e = (Yw - Y0 + ly + 8 * i) * kColTx + (Xw - X0 + lx);
                    part[2 * e] = acc[i].re;
                    part[2 * e + 1] = acc[i].im;
                }
                __syncthreads();  // s_item reuse
                return;
            }
            const int x = Xw + lx;
    #pragma unroll
            for (int i = 0; i < NG; i++) {
                const int y = Yw + ly + 8 * i;
                if (x >= g.N || y >= g.N) continue;
                const int64_t p = (int64_t)y * g.N + x;
                if (PLANAR)
                    flush_px(grid, grid_lo, p, NN + p, acc[i].re, acc[i].im, atomic, fresh);
                else
                    flush_px(grid, grid_lo, 2 * p, 2 * p + 1, acc[i].re, acc[i].im, atomic, fresh);
            }
            __syncthreads();  // s_item reuse
        };
        if constexpr (SC == kScBoth) {
            if (g.S == 64)
                body(std::integral_constant<int, 64>{});
            else
                body(std::integral_constant<int, 32>{});
        } else {
            body(std::integral_constant<int, SC>{});
        }
    }
}

// ---- small supports: one tap per lane --------------------------------------------
// At S <= 16 a wave's 16 x 16 block is mostly outside a visibility's S x S
// window (at S = 8 about one executed tap in eight is real), and the
// per-sub-kernel bookkeeping of 16 waves outweighs the few visibilities each
// sees.  Here lane l owns taps l, l + 64, ... of the window (row t / S,
// column t % S) and holds their GCF values for the current sub-kernel (the
// next sub-kernel's are loaded meanwhile); a wave takes one visibility at a
// time and adds its products into the workgroup's T x T tile in LDS with
// LDS atomics.  The tile holds doubles for every arithmetic form: on gfx950
// ds_add_f64 is cheap where ds_add_f32 is not, and the sums are the more
// accurate for it (planar re / im planes, row pitch 8 mod 32 words, so the
// rows of an 8- or 16-wide window spread over the banks).  An item's bin
// bounds for all its sub-kernels are read into LDS at once.  A sub-kernel
// with many visibilities in the tile is shared by the waves round-robin; the
// sparse ones are dealt out whole, one wave each.  The
// tile is flushed as the owner kernel flushes it.  LDS atomics add in no
// fixed order, so this kernel serves the atomic TILED strategy.
constexpr int kTapMaxS = 16;  // at most 4 taps per lane
constexpr int kTapWaves = 16;  // a power of two (round-robin by mask)
constexpr int kTapShare = 64;  // visibilities of a sub-kernel in a tile from which all waves share it
constexpr int kTapRows = 5;   // bin rows per tile: (T + S - 2) / 16 + 1 <= 5 at T <= 64, S <= 16
__host__ __device__ constexpr int tap_pitch(int T) { return T + 4; }  // doubles

// two workgroups per CU: 8 waves per SIMD, at most 64 VGPRs.  GM: 0 table
// GCF, 1 constant tap (COMPUTE_GCF and the coverage map: every tap 1, the
// record carries the transformed value), 2 analytic GCF.  SC: every job's
// support as a compile-time constant (0: read per job).
template <typename G, typename R, typename TO, bool PLANAR, int GM, int SC>
__global__ __launch_bounds__(64 * kTapWaves) __attribute__((amdgpu_waves_per_eu(8))) void k_taps(const JobDesc* __restrict__ jobs,
                                                         const Item* __restrict__ items,
                                                         QueueState* __restrict__ qs)
{
    using P = typename Stage<G>::P;  // product type
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* tacc = (double*)smem_raw;
    __shared__ int s_item;
    __shared__ int s_bnd[kSubKernels * kTapRows * 2];  // per sub-kernel and bin row: record range
    __shared__ Cplx<double> s_fac[2][kOversample][kTapMaxS];  // analytic factors (GM == 2)

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform

    for (;;) {
        if (tid == 0) s_item = atomicAdd(&qs->head, 1);
        __syncthreads();
        const int it_idx = s_item;
        if (it_idx >= qs->nitems) break;
        const Item it = items[it_idx];
        const int tile = __builtin_amdgcn_readfirstlane(it.tile);
        const int klo = __builtin_amdgcn_readfirstlane(it.klo);
        const int khi = __builtin_amdgcn_readfirstlane(it.khi);
        const int job = __builtin_amdgcn_readfirstlane(it.job);
        const TileGeom g = jobs[job].g;
        ConstPtr<R> rec = (ConstPtr<R>)(uintptr_t)jobs[job].rec;
        GlobalPtr<int32_t> offs = (GlobalPtr<int32_t>)(uintptr_t)jobs[job].offs;
        GlobalPtr<G> gtab = (GlobalPtr<G>)(uintptr_t)jobs[job].gcf;  // table in the staged type
        TO* __restrict__ grid = (TO*)jobs[job].grid;
        TO* __restrict__ grid_lo = (TO*)jobs[job].grid_lo;
        const bool atomic = jobs[job].split[tile] != 0;
        const bool fresh = jobs[job].fresh != 0;
        const int S = SC ? SC : g.S, T = g.T, SS = S * S;
        const int pitch = tap_pitch(T);
        const int plane = T * pitch;
        const int tx = tile % g.ntx, ty = tile / g.ntx;
        const int X0 = tx * T, Y0 = ty * T;
        // the bins of every visibility that reaches the tile (q in [X0, X0 + T + S - 2])
        const int bxl = X0 / g.H, bxh = min(g.nb - 1, (X0 + T + S - 2) / g.H);
        const int byl = Y0 / g.H, byh = min(g.nb - 1, (Y0 + T + S - 2) / g.H);
        const int nrow = byh - byl + 1;
        for (int e = tid; e < 2 * plane; e += blockDim.x) tacc[e] = 0.0;
        for (int e = tid; e < (khi - klo + 1) * nrow * 2; e += blockDim.x) {
            const int kk = klo + e / (nrow * 2), t = (e >> 1) % nrow;
            const int64_t key = ((int64_t)kk * g.nb + byl + t) * g.nb;
            s_bnd[(kk * kTapRows + t) * 2 + (e & 1)] = offs[(e & 1) ? key + bxh + 1 : key + bxl];
        }
        // per tap: row, column and LDS offset relative to the window's corner
        const int ntap = (SS + 63) / 64;
        int tr[4], tc[4], to[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            tr[i] = (lane + 64 * i) / S;
            tc[i] = (lane + 64 * i) % S;
            to[i] = tr[i] * pitch + tc[i];
        }
        // Sub-kernels with many visibilities in the tile are shared by all
        // waves (round-robin by visibility); the others are dealt out whole,
        // one wave each, so their bookkeeping is not repeated by every wave.
        uint64_t nz_mask, big_mask;
        {
            const int cnt =
                (lane >= klo && lane <= khi) ? jobs[job].tcount[(int64_t)tile * kSubKernels + lane] : 0;
            const uint64_t small_all = __ballot(cnt > 0 && cnt < kTapShare);
            big_mask = __ballot(cnt >= kTapShare);
            const int rank = __builtin_popcountll(small_all & ((1ull << lane) - 1));
            const uint64_t mine = __ballot(((small_all >> lane) & 1) && (rank & (kTapWaves - 1)) == w);
            nz_mask = big_mask | mine;
        }
        if (GM == 2) {
            // separable analytic GCF, as the owner kernel stages it: G[r][c] =
            // f_y(r) f_x(c), f(d) = exp(-d^2/(2 sigma^2)) exp(i beta d^2) with d
            // the tap's offset from the visibility for each sub-pixel position
            const double sigma = jobs[job].sigma, beta = jobs[job].beta;
            for (int e = tid; e < 2 * kOversample * S; e += blockDim.x) {
                const int axis = e / (kOversample * S), sp = (e / S) % kOversample, t = e % S;
                const double d = (double)(t - S / 2 + 1) - sp / 8.0;
                const double q = d * d;
                const double amp = exp(-q / (2.0 * sigma * sigma));
                double sn, cs;
                sincos(beta * q, &sn, &cs);
                s_fac[axis][sp][t].re = amp * cs;
                s_fac[axis][sp][t].im = amp * sn;
            }
        }
        auto load_taps = [&](int kk, P(&gr)[4], P(&gi)[4]) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                gr[i] = gi[i] = (P)0;
                if (i < ntap && lane + 64 * i < SS) {
                    if (GM == 1) {
                        gr[i] = (P)1;
                    } else if (GM == 2) {
                        const Cplx<double> a = s_fac[1][kk / kOversample][tr[i]];
                        const Cplx<double> b = s_fac[0][kk % kOversample][tc[i]];
                        Stage<G>::get(Stage<G>::make(a.re * b.re - a.im * b.im,
                                                     a.re * b.im + a.im * b.re),
                                      gr[i], gi[i]);
                    } else {
                        Stage<G>::get(gload<G>(gtab, (int64_t)kk * SS + lane + 64 * i), gr[i],
                                      gi[i]);
                    }
                }
            }
        };
        P gr[4], gi[4];
        if (GM != 2 && nz_mask) load_taps(__builtin_ctzll(nz_mask), gr, gi);
        __syncthreads();  // tile cleared, bounds listed (and the analytic factors)
        if (GM == 2 && nz_mask) load_taps(__builtin_ctzll(nz_mask), gr, gi);
        int handed = 0;   // visibilities of the item handed out so far, mod kTapWaves
        for (uint64_t todo = nz_mask; todo; todo &= todo - 1) {
            const int k = __builtin_ctzll(todo);
            const uint64_t rest = todo & (todo - 1);  // the sub-kernels after k
            P ngr[4], ngi[4];
            if (rest) load_taps(__builtin_ctzll(rest), ngr, ngi);
            const bool shared = (big_mask >> k) & 1;
            // lanes 2t and 2t + 1 hold the record range of bin row t
            const int bv = lane < 2 * nrow ? s_bnd[k * kTapRows * 2 + lane] : 0;
#pragma unroll
            for (int t = 0; t < kTapRows; t++) {
                if (t >= nrow) break;
                const int r0 = __builtin_amdgcn_readlane(bv, 2 * t);
                const int r1 = __builtin_amdgcn_readlane(bv, 2 * t + 1);
                // a shared sub-kernel: this wave's visibilities have running
                // index = w (mod kTapWaves); a dealt one: all of them
                const int i0 = shared ? (w - handed) & (kTapWaves - 1) : 0;
                const int step = shared ? kTapWaves : 1;
                for (int r = r0 + i0; r < r1; r += step) {
                    const uint32_t q = rec[r].q;
                    const P vr = (P)rec[r].vr, vi = (P)rec[r].vi;
                    // tap (row, column) lands on pixel q - (S - 1) + (column, row)
                    const int sx = (int)(q & 0xffffu) - (S - 1) - X0;
                    const int sy = (int)(q >> 16) - (S - 1) - Y0;
                    const int se = sy * pitch + sx;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        if (i >= ntap) break;
                        const int x = sx + tc[i], y = sy + tr[i];
                        if (lane + 64 * i < SS && (unsigned)x < (unsigned)T &&
                            (unsigned)y < (unsigned)T) {
                            const int e = se + to[i];
                            atomicAdd(&tacc[e], (double)(vr * gr[i] - vi * gi[i]));
                            atomicAdd(&tacc[plane + e], (double)(vr * gi[i] + vi * gr[i]));
                        }
                    }
                }
                if (shared) handed = (handed + (r1 - r0)) & (kTapWaves - 1);
            }
            if (rest) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    gr[i] = ngr[i];
                    gi[i] = ngi[i];
                }
            }
        }
        __syncthreads();
        const int64_t NN = (int64_t)g.N * g.N;
        for (int e = tid; e < T * T; e += blockDim.x) {
            const int ey = e / T, ex = e % T;
            const int x = X0 + ex, y = Y0 + ey;
            if (x >= g.N || y >= g.N) continue;
            const int le = ey * pitch + ex;
            const int64_t p = (int64_t)y * g.N + x;
            if (PLANAR)
                flush_px(grid, grid_lo, p, NN + p, tacc[le], tacc[plane + le], atomic, fresh);
            else
                flush_px(grid, grid_lo, 2 * p, 2 * p + 1, tacc[le], tacc[plane + le], atomic, fresh);
        }
        __syncthreads();  // the tile, the bounds and s_item are reused
    }
}

// Ordered reduction of split tiles (gather / window): each listed tile's
// partial tiles are summed per pixel in item order and flushed once -- a
// plain store into a fresh grid, else read-add-write (the tile is this
// block's alone).  The grid is the same bit for bit on every run.
template <typename TA, typename TO>
__global__ __launch_bounds__(256) void k_reduce_tiles(const JobDesc* __restrict__ jobs,
                                                      const QueueState* __restrict__ qs,
                                                      int planar)
{
    const int nr = min(qs->nreduce, jobs[0].rcap);
    const ReduceItem* __restrict__ list = jobs[0].reduce;
    for (int r = blockIdx.x; r < nr; r += gridDim.x) {
        const ReduceItem ri = list[r];
        if (ri.n <= 0) continue;
        const JobDesc& d = jobs[ri.job];
        const TileGeom g = d.g;
        const int tx = ri.tile % g.ntx, ty = ri.tile / g.ntx;
        const int64_t slot = (int64_t)g.T * g.Ty * 2;
        const TA* __restrict__ part = (const TA*)d.partials + (int64_t)ri.pslot * slot;
        TO* grid = (TO*)d.grid;
        TO* lo = (TO*)d.grid_lo;
        const bool fresh = d.fresh != 0;
        const int64_t NN = (int64_t)g.N * g.N;
        // a pixel's (re, im) partials by one 16-byte load per slot, four slots
        // in flight, added in slot order (the order fixes the result)
        using V2 = typename Pair2<TA>::T;
        const V2* __restrict__ p2 = (const V2*)part;
        const int64_t s2 = slot / 2;
        for (int e = threadIdx.x; e < g.T * g.Ty; e += blockDim.x) {
            const int x = tx * g.T + e % g.T, y = ty * g.Ty + e / g.T;
            if (x >= g.N || y >= g.N) continue;
            TA re = 0, im = 0;
            int i = 0;
            for (; i + 4 <= ri.n; i += 4) {
                V2 v[4];
#pragma unroll
                for (int u = 0; u < 4; u++) v[u] = p2[(i + u) * s2 + e];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    re += v[u].x;
                    im += v[u].y;
                }
            }
            for (; i < ri.n; i++) {
                const V2 v = p2[i * s2 + e];
                re += v.x;
                im += v.y;
            }
            const int64_t p = (int64_t)y * g.N + x;
            if (planar)
                flush_px(grid, lo, p, NN + p, re, im, false, fresh);
            else
                flush_px(grid, lo, 2 * p, 2 * p + 1, re, im, false, fresh);
        }
    }
}

// The GCF table in the staged type (FP32 or BF16 pairs), converted once per
// call so staging and the direct path move 8- or 4-byte elements.
template <typename TI, typename G>
__global__ __launch_bounds__(256) void k_convert_gcf(const ConvJob* __restrict__ jobs)
{
    const ConvJob c = jobs[blockIdx.y];
    const TI* __restrict__ src = (const TI*)c.src;
    G* __restrict__ dst = (G*)c.dst;
    // a table of support s_src embedded in one of s_dst: sub-kernel element
    // (r, c) moves to (r + o, c + o), o = ceil(s_dst / 2) - ceil(s_src / 2),
    // which keeps every tap on its pixel; the rest is zero
    const int sd = c.s_dst, ss = c.s_src;
    const int o = (sd - sd / 2) - (ss - ss / 2);
    const int64_t per = (int64_t)sd * sd;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < c.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t e = i;
        bool in = true;
        if (sd != ss) {
            const int64_t kk = i / per, rc = i - kk * per;
            const int r = (int)(rc / sd) - o, cc = (int)(rc % sd) - o;
            in = r >= 0 && r < ss && cc >= 0 && cc < ss;
            e = (kk * ss + r) * ss + cc;
        }
        dst[i] = in ? Stage<G>::make((double)src[2 * e], (double)src[2 * e + 1])
                    : Stage<G>::make(0.0, 0.0);
    }
}

int launch_convert(const ConvJob* table, int n, int64_t max_elems, int stage, bool in64,
                   hipStream_t s)
{
    if (n <= 0) return MIGRID_OK;
    const int64_t per = (max_elems + 255) / 256;
    const int bx = (int)(per < 256 ? (per < 1 ? 1 : per) : 256);
    const dim3 grid(bx, n);
    if (stage == kStageBf16) {
        if (in64)
            hipLaunchKernelGGL((k_convert_gcf<double, Bf16x2>), grid, dim3(256), 0, s, table);
        else
            hipLaunchKernelGGL((k_convert_gcf<float, Bf16x2>), grid, dim3(256), 0, s, table);
    } else if (stage == kStageF32) {
        if (in64)
            hipLaunchKernelGGL((k_convert_gcf<double, float2>), grid, dim3(256), 0, s, table);
        else
            hipLaunchKernelGGL((k_convert_gcf<float, float2>), grid, dim3(256), 0, s, table);
    } else {
        if (in64)
            hipLaunchKernelGGL((k_convert_gcf<double, double2>), grid, dim3(256), 0, s, table);
        else
            hipLaunchKernelGGL((k_convert_gcf<float, double2>), grid, dim3(256), 0, s, table);
    }
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

// ---- host side ------------------------------------------------------------------
template <int WX, int WY, typename G, typename TA, typename R, typename TI,
          typename TO, bool PLANAR, int GM, int SC = 0>
int run_owner(const OwnerLaunch& L, hipStream_t s)
{
    auto kern = k_owner<WX, WY, G, TA, R, TI, TO, PLANAR, GM, SC>;
    if (L.lds > kMaxLds)
        return fail(MIGRID_ENOTSUP, "tiled: the staged sub-kernel needs %zu bytes of LDS (at most %zu)",
                    L.lds, kMaxLds);
    MG_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)L.lds));
    int per_cu = 0;
    MG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * WX * WY, L.lds));
    if (per_cu < 1) return fail(MIGRID_ENOTSUP, "tiled: kernel does not fit (lds=%zu)", L.lds);
    int dev = 0, ncu = 0;
    MG_HIP(hipGetDevice(&dev));
    MG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (getenv("MIGRID_VERBOSE")) {
        hipFuncAttributes fa;
        MG_HIP(hipFuncGetAttributes(&fa, (const void*)kern));
        fprintf(stderr,
                "[migrid] owner launch: %d CUs x %d WG/CU (occupancy %d), %d threads, "
                "lds dyn %zu + static %zu, vgpr %d, sgpr %d, spill %zu\n",
                ncu, L.wg_per_cu > 0 ? L.wg_per_cu : per_cu, per_cu, 64 * WX * WY, L.lds,
                (size_t)fa.sharedSizeBytes, fa.numRegs, 0, (size_t)fa.localSizeBytes);
    }
    if (L.wg_per_cu > 0) per_cu = L.wg_per_cu;
    hipLaunchKernelGGL(kern, dim3(ncu * per_cu), dim3(64 * WX * WY), L.lds, s, L.jobs, L.items,
                       L.qs, L.smax, L.stage_min > 0 ? L.stage_min : kStageMinDefault);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

template <typename G, typename TA, typename R, typename TI, typename TO, bool PLANAR, int GM,
          int SC = 0>
int run_owner_col(const OwnerLaunch& L, hipStream_t s)
{
    auto kern = k_owner_col<G, TA, R, TI, TO, PLANAR, GM, SC>;
    if (L.lds > kMaxLds)
        return fail(MIGRID_ENOTSUP, "tiled: the staged sub-kernel needs %zu bytes of LDS (at most %zu)",
                    L.lds, kMaxLds);
    MG_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)L.lds));
    int per_cu = 0;
    MG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 1024, L.lds));
    if (per_cu < 1) return fail(MIGRID_ENOTSUP, "tiled: kernel does not fit (lds=%zu)", L.lds);
    int dev = 0, ncu = 0;
    MG_HIP(hipGetDevice(&dev));
    MG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (getenv("MIGRID_VERBOSE")) {
        hipFuncAttributes fa;
        MG_HIP(hipFuncGetAttributes(&fa, (const void*)kern));
        fprintf(stderr,
                "[migrid] column owner launch: %d CUs x %d WG/CU (occupancy %d), lds dyn %zu + "
                "static %zu, vgpr %d, spill %zu\n",
                ncu, L.wg_per_cu > 0 ? L.wg_per_cu : per_cu, per_cu, L.lds,
                (size_t)fa.sharedSizeBytes, fa.numRegs, (size_t)fa.localSizeBytes);
    }
    if (L.wg_per_cu > 0) per_cu = L.wg_per_cu;
    hipLaunchKernelGGL(kern, dim3(ncu * per_cu), dim3(1024), L.lds, s, L.jobs, L.items, L.qs,
                       L.smax);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

// Tile shapes: the tile option 64 runs the column-block kernel (128 x 64
// tiles, 16 waves of 8 x 64 pixels); 32 and 16 a workgroup of (T/16)^2 waves,
// each owning a 16 x 16 block.
template <typename G, typename TA, typename R, typename TI, typename TO, bool PLANAR, int GM>
int run_shape(int T, const OwnerLaunch& L, hipStream_t s)
{
    if (T == 64) {
        // every job at S = 64 or at S = 32: the support as a constant
        if (L.sc == 64) return run_owner_col<G, TA, R, TI, TO, PLANAR, GM, 64>(L, s);
        if (L.sc == 32) return run_owner_col<G, TA, R, TI, TO, PLANAR, GM, 32>(L, s);
        if (L.sc == kScBoth) return run_owner_col<G, TA, R, TI, TO, PLANAR, GM, kScBoth>(L, s);
        // table jobs always run embedded in 32 or 64 (tiled_plan); only the
        // computed forms keep other supports
        if constexpr (GM == 0)
            return fail(MIGRID_ENOTSUP, "tiled: internal: table jobs of mixed supports");
        else
            return run_owner_col<G, TA, R, TI, TO, PLANAR, GM>(L, s);
    }
    if (T == 32) return run_owner<2, 2, G, TA, R, TI, TO, PLANAR, GM>(L, s);
    if (T == 16) return run_owner<1, 1, G, TA, R, TI, TO, PLANAR, GM>(L, s);
    return fail(MIGRID_ENOTSUP, "tiled: tile edge %d (16, 32 or 64 supported)", T);
}

template <typename G, typename TA, typename R, typename TI, typename TO>
int run_modes(const KernelKey& k, const OwnerLaunch& L, hipStream_t s)
{
    if (k.planar) {
        if (k.gm == 0) return run_shape<G, TA, R, TI, TO, true, 0>(k.T, L, s);
        if (k.gm == 1) return run_shape<G, TA, R, TI, TO, true, 1>(k.T, L, s);
        return run_shape<G, TA, R, TI, TO, true, 2>(k.T, L, s);
    }
    if (k.gm == 0) return run_shape<G, TA, R, TI, TO, false, 0>(k.T, L, s);
    if (k.gm == 1) return run_shape<G, TA, R, TI, TO, false, 1>(k.T, L, s);
    return run_shape<G, TA, R, TI, TO, false, 2>(k.T, L, s);
}

template <typename G, typename R, typename TO, bool PLANAR, int GM, int SC>
int run_taps(const OwnerLaunch& L, hipStream_t s)
{
    auto kern = k_taps<G, R, TO, PLANAR, GM, SC>;
    if (L.lds > kMaxLds)
        return fail(MIGRID_ENOTSUP, "tiled: the LDS tile needs %zu bytes (at most %zu)", L.lds,
                    kMaxLds);
    MG_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)L.lds));
    int per_cu = 0;
    MG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * kTapWaves, L.lds));
    if (per_cu < 1) return fail(MIGRID_ENOTSUP, "tiled: tap kernel does not fit (lds=%zu)", L.lds);
    int dev = 0, ncu = 0;
    MG_HIP(hipGetDevice(&dev));
    MG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (getenv("MIGRID_VERBOSE")) {
        hipFuncAttributes fa;
        MG_HIP(hipFuncGetAttributes(&fa, (const void*)kern));
        fprintf(stderr,
                "[migrid] tap launch: %d CUs x %d WG/CU (occupancy %d), lds dyn %zu, vgpr %d, "
                "spill %zu\n",
                ncu, L.wg_per_cu > 0 ? L.wg_per_cu : per_cu, per_cu, L.lds, fa.numRegs,
                (size_t)fa.localSizeBytes);
    }
    if (L.wg_per_cu > 0) per_cu = std::min(per_cu, L.wg_per_cu);
    hipLaunchKernelGGL(kern, dim3(ncu * per_cu), dim3(64 * kTapWaves), L.lds, s, L.jobs, L.items,
                       L.qs);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

template <typename G, typename R, typename TO, bool PLANAR>
int run_taps_support(const KernelKey& k, const OwnerLaunch& L, hipStream_t s)
{
    if (k.gm == 1) return run_taps<G, R, TO, PLANAR, 1, 0>(L, s);  // constant tap
    if (k.gm == 2) return run_taps<G, R, TO, PLANAR, 2, 0>(L, s);  // analytic
    // a table GCF with every job at S = 8 or at S = 16: the support as a constant
    if (L.sc == 8) return run_taps<G, R, TO, PLANAR, 0, 8>(L, s);
    if (L.sc == 16) return run_taps<G, R, TO, PLANAR, 0, 16>(L, s);
    return run_taps<G, R, TO, PLANAR, 0, 0>(L, s);
}

template <typename G, typename R, typename TO>
int run_taps_layout(const KernelKey& k, const OwnerLaunch& L, hipStream_t s)
{
    return k.planar ? run_taps_support<G, R, TO, true>(k, L, s)
                    : run_taps_support<G, R, TO, false>(k, L, s);
}

template <typename TI, typename R>
void launch_keyrec(const PackJob* table, int n, int64_t max_npts, int vt, uint32_t* keys, R* rec0,
                   hipStream_t s)
{
    const int64_t per = (max_npts + 255) / 256;
    const int64_t cap = n < 8192 ? 8192 / n : 1;
    const int bx = (int)(per < 1 ? 1 : per < cap ? per : cap);
#define MG_KEYREC(VT)                                                                          \
    hipLaunchKernelGGL((k_keyrec<TI, R, VT>), dim3(bx, n), dim3(256), 0, s, table, keys, rec0)
    switch (vt) {
    case kValCompatS: MG_KEYREC(kValCompatS); break;
    case kValCompatI: MG_KEYREC(kValCompatI); break;
    case kValDebug: MG_KEYREC(kValDebug); break;
    default: MG_KEYREC(kValPlain);
    }
#undef MG_KEYREC
}

template <typename R>
void launch_gather(const PackJob* table, int n, int64_t max_npts, const int32_t* perm,
                   const R* rec0, R* rec, hipStream_t s)
{
    const int64_t per = (max_npts + 255) / 256;
    const int64_t cap = n < 8192 ? 8192 / n : 1;
    const int bx = (int)(per < 1 ? 1 : per < cap ? per : cap);
    hipLaunchKernelGGL((k_gather<R>), dim3(bx, n), dim3(256), 0, s, table, perm, rec0, rec);
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int dispatch_owner(const KernelKey& k, const OwnerLaunch& L, hipStream_t s)
{
    const bool in64 = k.in_type == MIGRID_F64, out64 = k.out_type == MIGRID_F64;
    switch (k.stage) {
    case kStageBf16:
        if (out64)
            return in64 ? run_modes<Bf16x2, float, RecF32, double, double>(k, L, s)
                        : run_modes<Bf16x2, float, RecF32, float, double>(k, L, s);
        return in64 ? run_modes<Bf16x2, float, RecF32, double, float>(k, L, s)
                    : run_modes<Bf16x2, float, RecF32, float, float>(k, L, s);
    case kStageF32:
        if (k.acc64)  // float products summed in double (native <double2,float2>)
            return run_modes<float2, double, RecF32, float, double>(k, L, s);
        if (out64)
            return in64 ? run_modes<float2, float, RecF32, double, double>(k, L, s)
                        : run_modes<float2, float, RecF32, float, double>(k, L, s);
        return in64 ? run_modes<float2, float, RecF32, double, float>(k, L, s)
                    : run_modes<float2, float, RecF32, float, float>(k, L, s);
    default:
        return run_modes<double2, double, RecF64, double, double>(k, L, s);
    }
}

int dispatch_taps(const KernelKey& k, const OwnerLaunch& L, hipStream_t s)
{
    const bool out64 = k.out_type == MIGRID_F64;
    switch (k.stage) {
    case kStageBf16:
        return out64 ? run_taps_layout<Bf16x2, RecF32, double>(k, L, s)
                     : run_taps_layout<Bf16x2, RecF32, float>(k, L, s);
    case kStageF32:  // float products (summed in double in the tile, as acc64 asks)
        return out64 ? run_taps_layout<float2, RecF32, double>(k, L, s)
                     : run_taps_layout<float2, RecF32, float>(k, L, s);
    default:
        return run_taps_layout<double2, RecF64, double>(k, L, s);
    }
}

TileCfg pick_tile(int S, int tile_opt, bool taps)
{
    TileCfg c;
    c.T = tile_opt > 0 ? tile_opt : 64;
    c.Ty = c.T;
    c.H = 16;  // the 16 x 16 block of a wave (tiles of 16, 32), the tap kernel
    if (c.T == 64 && !taps) {
        // the column kernel: 128 x 64 tiles of 8 x 64 wave blocks, bins of 8
        c.T = kColTx;
        c.Ty = kColTy;
        c.H = kColH;
    }
    return c;
}

int tiled_plan(const GridArgs* a, int n, const migrid_opts& o, TiledPlan& p)
{
    if (n < 1) return fail(MIGRID_EINVAL, "tiled plan with %d jobs", n);
    if (n > 65535) return fail(MIGRID_ENOTSUP, "a tiled plan takes at most 65535 jobs (%d)", n);
    const bool f32arith = o.arith == MIGRID_ARITH_F32 || o.arith == MIGRID_ARITH_F32_BF16GCF;
    if (o.strategy != MIGRID_TILED && o.arith != MIGRID_ARITH_NATIVE)
        return fail(MIGRID_ENOTSUP, "gather/window run native arithmetic only");
    if (o.tile != 0 && o.tile != 16 && o.tile != 32 && o.tile != 64)
        return fail(MIGRID_EINVAL, "tile=%d (0, 16, 32 or 64)", o.tile);
    p = TiledPlan();
    p.jobs.resize(n);
    size_t off = 0;
    int64_t item_cap = 0, item_cap_t = 0, item_cap_2 = 0;
    int smax_owner = 0, smin_owner = 1 << 30, smax_t = 0, smin_t = 1 << 30;
    // small supports: one tap per lane into an LDS tile (atomic strategy)
    const int tap_smax = o.tap_smax == 0 ? kTapMaxS : std::min(o.tap_smax, kTapMaxS);
    for (int j = 0; j < n; j++) {
        const GridArgs& x = a[j];
        if (x.S > 64) return fail(MIGRID_ENOTSUP, "tiled strategies support S <= 64 (S=%d)", x.S);
        if (x.N + x.S - 2 >= 65536) return fail(MIGRID_ENOTSUP, "tiled: N + S - 2 must be < 65536");
        if (x.debug && (f32arith || x.out_type != MIGRID_F64))
            return fail(MIGRID_ENOTSUP, "coverage mode needs native arithmetic into a double grid");
        if (x.grid_lo && x.out_type != MIGRID_F32)
            return fail(MIGRID_ENOTSUP, "a compensation grid needs a float grid");
        KernelKey k;
        k.stage = o.arith == MIGRID_ARITH_F32_BF16GCF ? kStageBf16
                  : (f32arith || x.in_type == MIGRID_F32) ? kStageF32
                                                         : kStageF64;
        k.acc64 = !f32arith && x.in_type == MIGRID_F32 && x.out_type == MIGRID_F64;
        k.in_type = x.in_type;
        k.out_type = x.out_type;
        k.planar = x.planar;
        k.gm = (x.debug || x.gcf_mode == MIGRID_GCF_COMPAT) ? 1
               : x.gcf_mode == MIGRID_GCF_ANALYTIC        ? 2
                                                          : 0;
        const int t_edge = o.tile > 0 ? o.tile : 64;
        const bool taps = o.strategy == MIGRID_TILED && x.S <= tap_smax &&
                          (t_edge + x.S - 2) / 16 + 1 <= kTapRows;
        const TileCfg tc = pick_tile(x.S, o.tile, taps);
        k.T = t_edge;  // the tile option; each job's geometry is in t.g
        // The column kernel has constant-support code for S = 32 and 64; a
        // table job of another support runs as the next of them, its table
        // embedded in zeros (converted copy), with the same pixels per tap:
        // the extra taps add exact zeros.  Measured at 4096^2 / 2M
        // clustered, this is 1.5-2x faster than the general code.
        const int sk = (tc.T == kColTx && k.gm == 0 && x.S != 32 && x.S != 64)
                           ? (x.S < 32 ? 32 : 64)
                           : x.S;
        if (x.N + sk - 2 >= 65536) return fail(MIGRID_ENOTSUP, "tiled: N + S - 2 must be < 65536");
        int vt = kValPlain;
        if (x.debug)
            vt = kValDebug;
        else if (x.gcf_mode == MIGRID_GCF_COMPAT)
            vt = o.strategy == MIGRID_TILED ? kValCompatS : kValCompatI;
        if (j == 0) {
            p.key = k;
            p.vt = vt;
        } else if (!(k == p.key) || vt != p.vt) {
            return fail(MIGRID_ENOTSUP,
                        "batched jobs must share types, layout, GCF mode, coverage mode and tile");
        }
        TiledJob& t = p.jobs[j];
        t.a = x;
        t.bg = bin_geom(x.N, sk, tc.H);
        t.g.N = x.N;
        t.g.S = sk;
        t.g.T = tc.T;
        t.g.Ty = tc.Ty;
        t.g.H = tc.H;
        t.g.nb = t.bg.nb;
        t.g.ntx = (x.N + tc.T - 1) / tc.T;
        t.g.nty = (x.N + tc.Ty - 1) / tc.Ty;
        t.rec64 = k.stage == kStageF64;
        t.vt = vt;
        t.e_base = p.E;
        t.k_base = p.K;
        p.E += x.npts;
        p.K += t.bg.nkeys + 1;  // the job's keys and its sentinel
        p.max_npts = std::max<int64_t>(p.max_npts, x.npts);
        const int64_t ntiles = (int64_t)t.g.ntx * t.g.nty;
        t.off_split = off;
        off += align256((size_t)ntiles);
        t.off_tcnt = off;
        off += align256((size_t)ntiles * kSubKernels * 4);
        // a table in the staged type unless the caller's table already is one
        t.convert = k.gm == 0 && (sk != x.S || !((k.stage == kStageF64 && x.in_type == MIGRID_F64) ||
                                                  (k.stage == kStageF32 && x.in_type == MIGRID_F32)));
        // jobs of one table (same source and support) share its conversion
        for (int i = 0; t.convert && i < j; i++) {
            const TiledJob& u = p.jobs[i];
            if (u.convert && !u.conv_dup && u.a.gcf == x.gcf && u.a.S == x.S && u.g.S == sk) {
                t.off_gtab = u.off_gtab;
                t.conv_dup = true;
                break;
            }
        }
        if (t.convert && !t.conv_dup) {
            const size_t gsz = k.stage == kStageBf16 ? 4 : k.stage == kStageF32 ? 8 : 16;
            t.off_gtab = off;
            off += align256((size_t)kSubKernels * sk * sk * gsz);
            p.nconv++;
            p.max_gcf = std::max<int64_t>(p.max_gcf, (int64_t)kSubKernels * sk * sk);
        }
        t.taps = taps;
        (t.taps ? item_cap_t : item_cap) += ntiles * kSubKernels;
        if (t.taps) {
            p.taps = true;
            smax_t = std::max(smax_t, x.S);
            smin_t = std::min(smin_t, x.S);
        } else {
            smax_owner = std::max(smax_owner, sk);
            smin_owner = std::min(smin_owner, sk);
        }
        p.smax = std::max(p.smax, sk);
    }
    p.owner = smax_owner > 0;
    p.smax_owner = smax_owner;
    p.s_common = smin_owner == smax_owner ? smax_owner : 0;
    p.s_common_t = smin_t == smax_t ? smax_t : 0;
    // Otherwise (analytic or constant-tap jobs of other supports) column-
    // kernel jobs of support 32 get a queue and a launch of their own, so
    // they keep the S = 32 instantiation (the general one was 35-40% slower
    // on an S = 64 job).
    // Column-kernel jobs all at 64 or 32 (every table job, its support
    // embedded): one launch of the instantiation that holds both constant
    // forms and picks one per item.
    if (p.owner && p.s_common == 0 && p.key.T == 64) {
        bool only = true;
        for (const TiledJob& t : p.jobs)
            if (!t.taps && t.g.S != 32 && t.g.S != 64) only = false;
        if (only) p.s_common = kScBoth;
    }
    if (p.owner && p.s_common == 0 && p.key.T == 64) {
        int hi = 0, lo = 1 << 30;
        bool has32 = false;
        for (const TiledJob& t : p.jobs) {
            if (t.taps) continue;
            if (t.g.S == 32) {
                has32 = true;
            } else {
                hi = std::max(hi, t.g.S);
                lo = std::min(lo, t.g.S);
            }
        }
        if (has32 && hi > 0) {
            for (TiledJob& t : p.jobs) {
                if (t.taps || t.g.S != 32) continue;
                t.q2 = true;
                const int64_t nt = (int64_t)t.g.ntx * t.g.nty * kSubKernels;
                item_cap -= nt;
                item_cap_2 += nt;
            }
            p.owner2 = true;
            p.s_common = hi == lo ? hi : 0;
        }
    }
    if (p.K >= ((int64_t)1 << 31) || p.E >= ((int64_t)1 << 31))
        return fail(MIGRID_ENOTSUP, "tiled: %lld keys / %lld visibilities in one plan (limit 2^31)",
                    (long long)p.K, (long long)p.E);
    const bool rec64 = p.key.stage == kStageF64;
    // the column kernel addresses records by unsigned 32-bit byte offsets
    if (p.E * (int64_t)(rec64 ? sizeof(RecF64) : sizeof(RecF32)) >= ((int64_t)1 << 32))
        return fail(MIGRID_ENOTSUP,
                    "tiled: %lld visibilities in one plan exceed 2^32 bytes of %d-byte records; "
                    "split the batch",
                    (long long)p.E, rec64 ? (int)sizeof(RecF64) : (int)sizeof(RecF32));
    p.off_perm = off;
    off += align256((size_t)(p.E > 0 ? p.E : 1) * 4);
    p.off_offs = off;
    off += align256((size_t)p.K * 4);
    p.off_rec = off;
    off += align256((size_t)(p.E > 0 ? p.E : 1) * (rec64 ? sizeof(RecF64) : sizeof(RecF32)));
    p.off_rec0 = off;  // the records in input order, before the gather
    off += align256((size_t)(p.E > 0 ? p.E : 1) * (rec64 ? sizeof(RecF64) : sizeof(RecF32)));
    p.off_pjobs = off;
    off += align256(sizeof(PackJob) * n);
    p.off_cjobs = off;
    off += align256(sizeof(ConvJob) * n);
    p.off_desc = off;
    off += align256(sizeof(JobDesc) * n);
    p.off_items = off;
    off += align256((size_t)(item_cap > 0 ? item_cap : 1) * sizeof(Item));
    p.off_items_t = off;
    off += align256((size_t)(item_cap_t > 0 ? item_cap_t : 1) * sizeof(Item));
    p.off_items_2 = off;
    off += align256((size_t)(item_cap_2 > 0 ? item_cap_2 : 1) * sizeof(Item));
    // the owner kernel's queue, the small-support kernel's, the second owner
    // queue's (support-32 jobs beside others)
    p.off_qs = off;
    p.off_qs_t = off + sizeof(QueueState);
    p.off_qs_2 = off + 2 * sizeof(QueueState);
    off += align256(3 * sizeof(QueueState));
    // jobs that write one grid flush atomically, into a grid cleared beforehand
    {
        std::vector<const void*> grids;
        for (const TiledJob& t : p.jobs) grids.push_back(t.a.grid);
        std::sort(grids.begin(), grids.end());
        p.shared = std::adjacent_find(grids.begin(), grids.end()) != grids.end();
    }
    p.fresh = !o.accumulate && !p.shared;
    // 8192: at 8192^2 / S=64 / 10M half the split tiles of 4096 (whose
    // flushes are atomics) and no slower; larger items leave dense tiles to
    // too few workgroups
    p.target = o.vis_per_wave > 0 ? o.vis_per_wave : 8192;
    // the tap kernel's items: smaller by default, so dense tiles spread wider
    p.target_t = o.vis_per_wave > 0 ? o.vis_per_wave : 2048;
    // gather / window: split tiles are summed by an ordered reduction, not
    // atomics.  Tile tx lists bins tx*T/H .. (tx*T + T + S - 2)/H (T is a
    // multiple of H), so along x a visibility's bin is listed by at most
    // floor((T + S - 2) / T) + 1 tiles, along y by floor((Ty + S - 2) / Ty)
    // + 1, and by m = their product in all: the per-tile counts sum to at
    // most m E.  A split tile holds c > target of
    // them and makes at most ceil(c / target) < 2c / target items, so
    // 2 m E / target + 1 partial tiles and m E / target + 1 split tiles bound
    // what the work-item pass lists.  When those slots would pass 4 GiB the
    // work items grow until they fit.
    p.det = o.strategy != MIGRID_TILED && !p.shared;
    if (p.det) {
        int64_t m = 1;
        for (const TiledJob& t : p.jobs) {
            const int64_t mx = (t.g.T + t.g.S - 2) / t.g.T + 1;
            const int64_t my = (t.g.Ty + t.g.S - 2) / t.g.Ty + 1;
            m = std::max<int64_t>(m, mx * my);
        }
        const size_t tsz = (p.key.stage == kStageF64 || p.key.acc64) ? 8 : 4;  // accumulator
        int64_t tpx = 0;  // pixels of the largest tile
        for (const TiledJob& t : p.jobs) tpx = std::max<int64_t>(tpx, (int64_t)t.g.T * t.g.Ty);
        const int64_t slot = tpx * 2 * (int64_t)tsz;
        const int64_t max_slots = ((int64_t)4 << 30) / slot;
        if (2 * m * p.E / p.target + 1 > max_slots)
            p.target = (2 * m * p.E + max_slots - 2) / (max_slots - 1);
        p.partial_cap = 2 * m * p.E / p.target + 1;
        p.reduce_cap = m * p.E / p.target + 1;
        p.off_partials = off;
        off += align256((size_t)p.partial_cap * slot);
        p.off_reduce = off;
        off += align256((size_t)p.reduce_cap * sizeof(ReduceItem));
    }
    p.arena_bytes = off;
    p.bin_bytes = bin_scratch_bytes(p.E, p.K);
    p.item_cap = item_cap;
    p.item_cap_t = item_cap_t;
    p.item_cap_2 = item_cap_2;
    p.stage_min = o.stage_min;
    p.wg_per_cu = o.wg_per_cu;
    const size_t gsz = p.key.stage == kStageBf16 ? 4 : p.key.stage == kStageF32 ? 8 : 16;
    // (the column kernel stages S = 32 sub-kernels in pairs: two layouts)
    p.lds = !p.owner ? 0
            : p.key.T == 64
                ? (size_t)std::max(col_rows(smax_owner) * col_pitch(smax_owner),
                                   smax_owner >= 32 ? 2 * col_rows(32) * col_pitch(32) : 0) * gsz
                : (size_t)stage_rows(smax_owner) * stage_pitch(smax_owner) * gsz;
    p.lds_taps = p.taps ? (size_t)2 * p.key.T * tap_pitch(p.key.T) * sizeof(double) : 0;
    return MIGRID_OK;
}

void tiled_descs(const TiledPlan& p, char* arena, std::vector<JobDesc>& d)
{
    d.resize(p.jobs.size());
    for (size_t j = 0; j < p.jobs.size(); j++) {
        const TiledJob& t = p.jobs[j];
        d[j].g = t.g;
        d[j].rec = arena + p.off_rec;  // CSR offsets are positions in the shared records
        d[j].offs = (const int32_t*)(arena + p.off_offs) + t.k_base;
        d[j].tcount = (const int32_t*)(arena + t.off_tcnt);
        d[j].split = (const uint8_t*)(arena + t.off_split);
        d[j].gcf = t.convert ? (const void*)(arena + t.off_gtab) : t.a.gcf;
        d[j].grid = t.a.grid;
        d[j].grid_lo = t.a.grid_lo;
        d[j].sigma = t.a.sigma;
        d[j].beta = t.a.beta;
        d[j].fresh = p.fresh ? 1 : 0;
        d[j].shared = p.shared ? 1 : 0;
        d[j].det = p.det ? 1 : 0;
        d[j].taps = t.taps ? 1 : 0;
        d[j].q2 = t.q2 ? 1 : 0;
        d[j].pcap = (int32_t)p.partial_cap;
        d[j].rcap = (int32_t)p.reduce_cap;
        d[j].partials = p.det ? (void*)(arena + p.off_partials) : nullptr;
        d[j].reduce = p.det ? (ReduceItem*)(arena + p.off_reduce) : nullptr;
    }
}

int tiled_write_descs(const TiledPlan& p, char* arena, const std::vector<JobDesc>& d,
                      hipStream_t s)
{
    return put_table((JobDesc*)(arena + p.off_desc), d.data(), (int)d.size(), s);
}

// dst[i] = src[i] + delta: a kept sort moved between a job's own numbering
// and its place in a launch (delta = +/- the job's first element).
static __global__ __launch_bounds__(256) void k_rebase(int32_t* __restrict__ dst,
                                                const int32_t* __restrict__ src, int64_t n,
                                                int32_t delta)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i] + delta;
}

static int launch_rebase(int32_t* dst, const int32_t* src, int64_t n, int64_t delta, hipStream_t s)
{
    if (n <= 0) return MIGRID_OK;
    const int64_t b = (n + 255) / 256;
    hipLaunchKernelGGL(k_rebase, dim3((unsigned)(b < 8192 ? b : 8192)), dim3(256), 0, s, dst, src,
                       n, (int32_t)delta);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

size_t tiled_sort_bytes(const TiledJob& t)
{
    return align256((size_t)(t.a.npts > 0 ? t.a.npts : 1) * 4) + (size_t)(t.bg.nkeys + 1) * 4;
}

int tiled_keep_sort(const TiledPlan& p, const char* arena, int j, void* dst, hipStream_t s)
{
    const TiledJob& t = p.jobs[j];
    const int32_t* perm = (const int32_t*)(arena + p.off_perm) + t.e_base;
    const int32_t* offs = (const int32_t*)(arena + p.off_offs) + t.k_base;
    int32_t* d_perm = (int32_t*)dst;
    int32_t* d_offs = (int32_t*)((char*)dst + align256((size_t)(t.a.npts > 0 ? t.a.npts : 1) * 4));
    MG_TRY(launch_rebase(d_perm, perm, t.a.npts, -t.e_base, s));
    return launch_rebase(d_offs, offs, t.bg.nkeys + 1, -t.e_base, s);
}

int tiled_prepare(const TiledPlan& p, char* arena, void* bin_scratch, hipStream_t s, int n_sorted,
                  const void* const* kept)
{
    const int n = (int)p.jobs.size();
    if (n_sorted < 0 || n_sorted > n) n_sorted = n;
    QueueState* qs = (QueueState*)(arena + p.off_qs);
    Item* items = (Item*)(arena + p.off_items);
    int32_t* perm = (int32_t*)(arena + p.off_perm);
    int32_t* offs = (int32_t*)(arena + p.off_offs);
    MG_TRY(fill_u32(qs, 0u, 3 * sizeof(QueueState) / 4, s));  // the three queues
    // job tables
    std::vector<PackJob> pj(n);
    std::vector<ConvJob> cj;
    for (int j = 0; j < n; j++) {
        const TiledJob& t = p.jobs[j];
        pj[j].uv = t.a.uv;
        pj[j].vals = t.a.vals;
        pj[j].gcf = t.a.gcf;
        pj[j].offs = offs + t.k_base;
        pj[j].e_base = t.e_base;
        pj[j].npts = t.a.npts;
        pj[j].k_base = t.k_base;
        pj[j].nkeys = t.bg.nkeys;
        pj[j].N = t.bg.N;
        pj[j].S = t.g.S;
        pj[j].H = t.bg.H;
        pj[j].nb = t.bg.nb;
        pj[j].pitch = t.g.T == kColTx ? col_pitch(t.g.S) : stage_pitch(t.g.S);
        pj[j].esize = p.key.stage == kStageF64 ? 16 : p.key.stage == kStageF32 ? 8 : 4;
        if (t.convert && !t.conv_dup) {
            ConvJob c;
            c.src = t.a.gcf;
            c.dst = arena + t.off_gtab;
            c.n = (int64_t)kSubKernels * t.g.S * t.g.S;
            c.s_src = t.a.S;
            c.s_dst = t.g.S;
            cj.push_back(c);
        }
    }
    PackJob* dpj = (PackJob*)(arena + p.off_pjobs);
    MG_TRY(put_table(dpj, pj.data(), n, s));
    if (!cj.empty()) {
        ConvJob* dcj = (ConvJob*)(arena + p.off_cjobs);
        MG_TRY(put_table(dcj, cj.data(), (int)cj.size(), s));
        MG_TRY(launch_convert(dcj, (int)cj.size(), p.max_gcf, p.key.stage,
                              p.key.in_type == MIGRID_F64, s));
    }
    // every job's keys (into the sort's key buffer) and records, in input order
    void* rec = arena + p.off_rec;
    void* rec0 = arena + p.off_rec0;
    uint32_t* keys = sort_key_buffer(bin_scratch);
    if (p.E > 0) {
        if (p.key.in_type == MIGRID_F64) {
            if (p.key.stage == kStageF64)
                launch_keyrec<double, RecF64>(dpj, n, p.max_npts, p.vt, keys, (RecF64*)rec0, s);
            else
                launch_keyrec<double, RecF32>(dpj, n, p.max_npts, p.vt, keys, (RecF32*)rec0, s);
        } else {
            launch_keyrec<float, RecF32>(dpj, n, p.max_npts, p.vt, keys, (RecF32*)rec0, s);
        }
        MG_HIP(hipGetLastError());
    }
    // the jobs sorted here come first (their keys are the buffer's first
    // elements); the others' kept sorts are moved to their places in this launch
    if (n_sorted == n) {
        MG_TRY(sort_keys(p.E, p.K, bin_scratch, perm, offs, s));
    } else {
        if (n_sorted > 0) {
            const TiledJob& first = p.jobs[n_sorted];
            MG_TRY(sort_keys(first.e_base, first.k_base, bin_scratch, perm, offs, s));
        }
        for (int j = n_sorted; j < n; j++) {
            const TiledJob& t = p.jobs[j];
            const int32_t* k_perm = (const int32_t*)kept[j - n_sorted];
            const int32_t* k_offs = (const int32_t*)((const char*)k_perm +
                                                     align256((size_t)(t.a.npts > 0 ? t.a.npts : 1) * 4));
            MG_TRY(launch_rebase(perm + t.e_base, k_perm, t.a.npts, t.e_base, s));
            MG_TRY(launch_rebase(offs + t.k_base, k_offs, t.bg.nkeys + 1, t.e_base, s));
        }
    }
    if (p.E > 0) {
        if (p.key.stage == kStageF64)
            launch_gather<RecF64>(dpj, n, p.max_npts, perm, (const RecF64*)rec0, (RecF64*)rec, s);
        else
            launch_gather<RecF32>(dpj, n, p.max_npts, perm, (const RecF32*)rec0, (RecF32*)rec, s);
        MG_HIP(hipGetLastError());
    }
    int64_t max_tiles = 0, max_ntx = 0;
    for (const TiledJob& t : p.jobs) {
        max_tiles = std::max<int64_t>(max_tiles, (int64_t)t.g.ntx * t.g.nty);
        max_ntx = std::max<int64_t>(max_ntx, std::max(t.g.ntx, t.g.nty));
    }
    const JobDesc* ddesc = (const JobDesc*)(arena + p.off_desc);
    const int cb = (int)std::min<int64_t>(max_ntx * kSubKernels, n < 8 ? 8192 : 65536 / n);
    hipLaunchKernelGGL(k_tile_counts, dim3(cb, n), dim3(128), 0, s, ddesc);
    const dim3 ig((unsigned)((max_tiles + 3) / 4), n);
    Item* items_t = (Item*)(arena + p.off_items_t);
    QueueState* qs_t = (QueueState*)(arena + p.off_qs_t);
    Item* items_2 = (Item*)(arena + p.off_items_2);
    QueueState* qs_2 = (QueueState*)(arena + p.off_qs_2);
    if (p.key.out_type == MIGRID_F64)
        hipLaunchKernelGGL(k_items<double>, ig, dim3(256), 0, s, ddesc, p.target, p.target_t,
                           p.key.planar ? 1 : 0, items, qs, items_t, qs_t, items_2, qs_2);
    else
        hipLaunchKernelGGL(k_items<float>, ig, dim3(256), 0, s, ddesc, p.target, p.target_t,
                           p.key.planar ? 1 : 0, items, qs, items_t, qs_t, items_2, qs_2);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

int tiled_execute(const TiledPlan& p, char* arena, hipStream_t s)
{
    OwnerLaunch L;
    L.jobs = (const JobDesc*)(arena + p.off_desc);
    L.items = (const Item*)(arena + p.off_items);
    L.qs = (QueueState*)(arena + p.off_qs);
    L.smax = p.smax_owner;  // the owner kernel's LDS layout (p.lds) is sized for it
    L.lds = p.lds;
    L.stage_min = p.stage_min;
    L.wg_per_cu = p.wg_per_cu;
    if (p.taps) {
        OwnerLaunch Lt = L;
        Lt.items = (const Item*)(arena + p.off_items_t);
        Lt.qs = (QueueState*)(arena + p.off_qs_t);
        Lt.lds = p.lds_taps;
        Lt.sc = p.s_common_t;
        MG_TRY(dispatch_taps(p.key, Lt, s));
    }
    L.sc = p.s_common;
    if (p.owner) MG_TRY(dispatch_owner(p.key, L, s));
    if (p.owner2) {
        OwnerLaunch L2 = L;
        L2.items = (const Item*)(arena + p.off_items_2);
        L2.qs = (QueueState*)(arena + p.off_qs_2);
        L2.sc = 32;
        MG_TRY(dispatch_owner(p.key, L2, s));
    }
    if (p.det) {  // the split tiles, summed in item order
        const int nb = (int)std::min<int64_t>(p.reduce_cap, 2048);
        const int planar = p.key.planar ? 1 : 0;
        const bool acc_d = p.key.stage == kStageF64 || p.key.acc64;
        if (acc_d && p.key.out_type == MIGRID_F64)
            hipLaunchKernelGGL((k_reduce_tiles<double, double>), dim3(nb), dim3(256), 0, s, L.jobs,
                               L.qs, planar);
        else if (acc_d)
            hipLaunchKernelGGL((k_reduce_tiles<double, float>), dim3(nb), dim3(256), 0, s, L.jobs,
                               L.qs, planar);
        else if (p.key.out_type == MIGRID_F64)
            hipLaunchKernelGGL((k_reduce_tiles<float, double>), dim3(nb), dim3(256), 0, s, L.jobs,
                               L.qs, planar);
        else
            hipLaunchKernelGGL((k_reduce_tiles<float, float>), dim3(nb), dim3(256), 0, s, L.jobs,
                               L.qs, planar);
        MG_HIP(hipGetLastError());
    }
    return MIGRID_OK;
}

// One job through the tiled strategies: the plan, then (unless there is
// nothing to grid) its passes on `arena` (p.arena_bytes) and `bins`
// (p.bin_bytes), which the caller provides.
static int run_single(const TiledPlan& p, const GridArgs& a, char* arena, void* bins,
                      Profiler* prof, hipStream_t s)
{
    std::vector<JobDesc> d;
    tiled_descs(p, arena, d);
    MG_TRY(tiled_write_descs(p, arena, d, s));
    MG_TRY(tiled_prepare(p, arena, bins, s));
    if (prof) MG_TRY(prof->begin(s));
    MG_TRY(tiled_execute(p, arena, s));
    if (prof) MG_TRY(prof->end(s));
    return MIGRID_OK;
}

// Nothing to grid: a fresh grid is only cleared (fill kernels).
static int clear_only(const TiledPlan& p, const GridArgs& a, hipStream_t s)
{
    if (!p.fresh) return MIGRID_OK;
    const size_t gb = (size_t)a.N * a.N * 2 * (a.out_type == MIGRID_F64 ? 8 : 4);
    MG_TRY(launch_zero(a.grid, gb, s));
    if (a.grid_lo) MG_TRY(launch_zero(a.grid_lo, gb, s));
    return MIGRID_OK;
}

// Visibilities one plan of this job's form takes: its packed records stay
// under 2^32 bytes (32-byte records with FP64 tiles, 16-byte otherwise).
static int64_t tiled_chunk_limit(const GridArgs& a, const migrid_opts& o)
{
    const bool f32arith = o.arith == MIGRID_ARITH_F32 || o.arith == MIGRID_ARITH_F32_BF16GCF;
    const bool rec64 = o.arith != MIGRID_ARITH_F32_BF16GCF && !f32arith && a.in_type != MIGRID_F32;
    return (((int64_t)1 << 32) - 1) / (int64_t)(rec64 ? sizeof(RecF64) : sizeof(RecF32));
}

int launch_tiled(migrid_ctx* ctx, const GridArgs& a, const migrid_opts& o, hipStream_t s)
{
    if (!ctx) return fail(MIGRID_EINVAL, "tiled strategies need a context");
    // a job past one plan's records runs as consecutive chunks of its
    // visibilities, the first into the grid as the caller asked (fresh or
    // accumulating), the rest added to it; the chunk order is fixed, so the
    // ordered-reduction strategies stay reproducible
    const int64_t lim = tiled_chunk_limit(a, o);
    const size_t esz = a.in_type == MIGRID_F64 ? 8 : 4;
    for (int64_t done = 0;;) {
        GridArgs c = a;
        migrid_opts oc = o;
        c.npts = std::min<int64_t>(a.npts - done, lim);
        c.uv = (const char*)a.uv + done * 2 * esz;
        c.vals = (const char*)a.vals + done * 2 * esz;
        if (done > 0) oc.accumulate = 1;
        TiledPlan p;
        MG_TRY(tiled_plan(&c, 1, oc, p));
        if (c.npts == 0) return clear_only(p, c, s);
        MG_TRY(ctx->bins.ensure(p.bin_bytes));
        MG_TRY(ctx->records.ensure(p.arena_bytes));
        MG_TRY(run_single(p, c, (char*)ctx->records.p, ctx->bins.p, &ctx->prof, s));
        done += c.npts;
        if (done >= a.npts) return MIGRID_OK;
    }
}

int launch_tiled_scratch(const GridArgs& a, const migrid_opts& o, void* scratch, size_t cap,
                         hipStream_t s)
{
    TiledPlan p;
    MG_TRY(tiled_plan(&a, 1, o, p));
    if (a.npts == 0) return clear_only(p, a, s);
    const size_t need = align256(p.arena_bytes) + p.bin_bytes;
    if (need > cap)
        return fail(MIGRID_ENOTSUP,
                    "this job needs %zu bytes of device scratch, beyond the %zu bytes of the "
                    "allocation-free entry point (grid it through a context instead)",
                    need, cap);
    char* arena = (char*)scratch;
    return run_single(p, a, arena, arena + align256(p.arena_bytes), nullptr, s);
}

}  // namespace migrid
