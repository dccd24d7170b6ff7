// Bin sort of visibilities by (sub-kernel k, bin_y, bin_x).
//
// Replaces the reference's host qsort (grid.cu:467-475), vis2ints
// (grid_gpu.cu:352-365) and set_bookmarks (:366-390) with an on-device,
// stable sort:
//   1. k_keys     per visibility: q = floor(uv) + S/2, k = 8*sub_y + sub_x,
//                 key = (k*nb + q_y/H)*nb + q_x/H (or the sentinel nkeys for a
//                 visibility that touches no pixel) and a key histogram;
//   2. LSD radix sort of (key, index) in 8-bit digits; each pass ranks a block's
//      2048 elements stably with wave64 ballots (a lane's peers are the lanes
//      whose digit matches in all 8 bit-ballots; its rank is the popcount of
//      the peers below it, __builtin_amdgcn_mbcnt) and an LDS prefix over the
//      block's waves;
//   3. CSR offsets = exclusive scan of the key histogram (the bookmarks).
// The permutation is the stable order, so it is bit-identical to a serial
// counting sort of the same keys.
#include <hip/hip_runtime.h>

#include "bin.hpp"
#include "device_common.hpp"
#include "internal.hpp"

namespace migrid {
namespace {

constexpr int kRsThreads = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsThreads * kRsItems;  // elements per block
constexpr int kScanItems = 8;
constexpr int kScanTile = 256 * kScanItems;

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const int lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

__device__ __forceinline__ int mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

template <typename TI>
__global__ __launch_bounds__(256) void k_keys(const TI* __restrict__ uv, int64_t npts, int N, int S,
                                              int H, int nb, uint32_t* __restrict__ keys,
                                              uint32_t* __restrict__ idx,
                                              int32_t* __restrict__ hist)
{
    const uint32_t nkeys = (uint32_t)kSubKernels * nb * nb;
    const int qmax = N + S - 2;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < npts;
         n += (int64_t)gridDim.x * blockDim.x) {
        const TI u = uv[2 * n], v = uv[2 * n + 1];
        const TI fu = floor_t(u), fv = floor_t(v);
        const int sx = (int)floor_t((TI)kOversample * (u - fu));
        const int sy = (int)floor_t((TI)kOversample * (v - fv));
        // q computed in the input type's range before narrowing: a coordinate
        // far outside the grid must not wrap into it
        const TI qxf = fu + (TI)(S / 2), qyf = fv + (TI)(S / 2);
        uint32_t key = nkeys;
        if (qxf >= 0 && qyf >= 0 && qxf <= (TI)qmax && qyf <= (TI)qmax) {
            const int qx = (int)qxf, qy = (int)qyf;
            key = ((uint32_t)(kOversample * sy + sx) * nb + qy / H) * nb + qx / H;
            atomicAdd(&hist[key], 1);
        }
        keys[n] = key;
        idx[n] = (uint32_t)n;
    }
}

// Per-block digit counts, stored digit-major: counts[d * nblk + blk].
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t* __restrict__ keys, int64_t n,
                                                        int shift, int32_t* __restrict__ counts)
{
    __shared__ int h[256];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int j = 0; j < kRsItems; j++) {
        const int64_t i = base + j * kRsThreads + t;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    counts[(int64_t)t * gridDim.x + blockIdx.x] = h[t];
}

// Stable scatter of one radix pass.  Elements are visited round-major
// (round j, thread t -> tile position j*256 + t), which is index order, and
// each round is ranked with ballots plus an LDS prefix over the 4 waves.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift, const int32_t* __restrict__ base_offs)
{
    __shared__ int wave_cnt[4][256];
    __shared__ int running[256];
    __shared__ int blk_base[256];
    const int t = threadIdx.x;
    const int w = t >> 6;
    running[t] = 0;
    blk_base[t] = base_offs[(int64_t)t * gridDim.x + blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int j = 0; j < kRsItems; j++) {
        for (int ww = 0; ww < 4; ww++) wave_cnt[ww][t] = 0;
        __syncthreads();
        const int64_t i = base + j * kRsThreads + t;
        const bool valid = i < n;
        uint32_t key = 0, val = 0;
        int d = 0;
        if (valid) {
            key = kin[i];
            val = vin[i];
            d = (key >> shift) & 255;
        }
        // peers: active lanes with the same digit
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            const uint64_t m = __ballot(valid && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        const int rank = mbcnt(peers);
        if (valid && (peers & lanemask_lt()) == 0) wave_cnt[w][d] = __popcll(peers);
        __syncthreads();
        {
            int acc = running[t];
            for (int ww = 0; ww < 4; ww++) {
                const int c = wave_cnt[ww][t];
                wave_cnt[ww][t] = acc;
                acc += c;
            }
            running[t] = acc;
        }
        __syncthreads();
        if (valid) {
            const int64_t pos = (int64_t)blk_base[d] + wave_cnt[w][d] + rank;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
    }
}

// ---- exclusive scan of int32 ------------------------------------------------
__device__ __forceinline__ int wave_incl_scan(int x)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                     int32_t* __restrict__ partial)
{
    __shared__ int ws[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int s = 0;
    for (int j = 0; j < kScanItems; j++) {
        const int64_t i = base + j * 256 + threadIdx.x;
        if (i < n) s += in[i];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// Single block: exclusive scan of the block partials in place.
__global__ __launch_bounds__(1024) void k_scan_partials(int32_t* __restrict__ p, int64_t n)
{
    __shared__ int ws[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t b = 0; b < n; b += 1024) {
        const int64_t i = b + threadIdx.x;
        const int x = i < n ? p[i] : 0;
        int inc = wave_incl_scan(x);
        if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = inc;
        __syncthreads();
        int wbase = 0;
        for (int w = 0; w < (int)(threadIdx.x >> 6); w++) wbase += ws[w];
        const int c = carry;
        if (i < n) p[i] = c + wbase + inc - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c + wbase + inc;
        __syncthreads();
    }
}

// out[i] = partial[blk] + exclusive prefix within the block; out[n] = total
// when write_total.  Each thread owns kScanItems consecutive elements.
__global__ __launch_bounds__(256) void k_scan_final(const int32_t* __restrict__ in, int64_t n,
                                                    const int32_t* __restrict__ partial,
                                                    int32_t* __restrict__ out, int write_total)
{
    __shared__ int ws[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
    for (int j = 0; j < kScanItems; j++) {
        const int64_t i = base + j;
        v[j] = i < n ? in[i] : 0;
        s += v[j];
    }
    const int inc = wave_incl_scan(s);
    if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = inc;
    __syncthreads();
    int run = partial[blockIdx.x] + inc - s;
    for (int w = 0; w < (int)(threadIdx.x >> 6); w++) run += ws[w];
    for (int j = 0; j < kScanItems; j++) {
        const int64_t i = base + j;
        if (i < n) out[i] = run;
        run += v[j];
        if (write_total && i == n - 1) out[n] = run;
    }
}

}  // namespace

int64_t scan_scratch_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

int exclusive_scan(const int32_t* in, int64_t n, int32_t* out, int32_t* partial, bool write_total,
                   hipStream_t s)
{
    if (n <= 0) {
        if (write_total) MG_TRY(fill_u32(out, 0u, 1, s));
        return MIGRID_OK;
    }
    const int64_t nblk = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nblk), dim3(256), 0, s, in, n, partial);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, partial, nblk);
    hipLaunchKernelGGL(k_scan_final, dim3(nblk), dim3(256), 0, s, in, n, partial, out,
                       write_total ? 1 : 0);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

BinGeom bin_geom(int64_t N, int64_t S, int H)
{
    BinGeom g;
    g.N = (int)N;
    g.S = (int)S;
    g.H = H;
    g.nb = (int)((N + S - 2) / H + 1);
    g.nkeys = (int64_t)kSubKernels * g.nb * g.nb;
    return g;
}

size_t bin_scratch_bytes(int64_t npts, const BinGeom& g)
{
    const int64_t nblk = (npts + kRsTile - 1) / kRsTile;
    size_t b = 0;
    b += 4 * (size_t)npts * 4;                         // keys/idx double buffers
    b += (size_t)(g.nkeys + 1) * 4;                    // key histogram
    b += (size_t)256 * (nblk + 1) * 4 * 2;             // digit counts + offsets
    const int64_t big = g.nkeys + 1 > 256 * (nblk + 1) ? g.nkeys + 1 : 256 * (nblk + 1);
    b += (size_t)scan_scratch_elems(big) * 4;          // scan partials
    return b + 64 * 5;
}

int bin_sort(const BinGeom& g, int in_type, const void* uv, int64_t npts, void* scratch,
             int32_t* perm, int32_t* offsets, hipStream_t s)
{
    const int64_t nblk = (npts + kRsTile - 1) / kRsTile;
    char* p = (char*)scratch;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += (bytes + 63) & ~(size_t)63;
        return r;
    };
    uint32_t* k0 = (uint32_t*)take((size_t)npts * 4);
    uint32_t* v0 = (uint32_t*)take((size_t)npts * 4);
    uint32_t* k1 = (uint32_t*)take((size_t)npts * 4);
    uint32_t* v1 = (uint32_t*)take((size_t)npts * 4);
    int32_t* hist = (int32_t*)take((size_t)(g.nkeys + 1) * 4);
    int32_t* counts = (int32_t*)take((size_t)256 * (nblk + 1) * 4);
    int32_t* coffs = (int32_t*)take((size_t)256 * (nblk + 1) * 4);
    const int64_t big = g.nkeys + 1 > 256 * (nblk + 1) ? g.nkeys + 1 : 256 * (nblk + 1);
    int32_t* partial = (int32_t*)take((size_t)scan_scratch_elems(big) * 4);

    MG_TRY(fill_u32(hist, 0u, g.nkeys + 1, s));
    if (npts > 0) {
        int blocks = (int)((npts + 255) / 256 < 4096 ? (npts + 255) / 256 : 4096);
        if (in_type == MIGRID_F64)
            hipLaunchKernelGGL(k_keys<double>, dim3(blocks), dim3(256), 0, s, (const double*)uv,
                               npts, g.N, g.S, g.H, g.nb, k0, v0, hist);
        else
            hipLaunchKernelGGL(k_keys<float>, dim3(blocks), dim3(256), 0, s, (const float*)uv, npts,
                               g.N, g.S, g.H, g.nb, k0, v0, hist);
        MG_HIP(hipGetLastError());
        // number of key bits (the sentinel nkeys included)
        int bits = 1;
        while (((int64_t)1 << bits) <= g.nkeys) bits++;
        int passes = (bits + 7) / 8;
        uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
        for (int pass = 0; pass < passes; pass++) {
            const int shift = 8 * pass;
            if (pass == passes - 1) vo = (uint32_t*)perm;  // last pass writes the permutation
            hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(kRsThreads), 0, s, ki, npts, shift,
                               counts);
            MG_TRY(exclusive_scan(counts, 256 * nblk, coffs, partial, false, s));
            hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(kRsThreads), 0, s, ki, vi, ko, vo,
                               npts, shift, coffs);
            MG_HIP(hipGetLastError());
            uint32_t* tk = ki;
            ki = ko;
            ko = tk;
            uint32_t* tv = vi;
            vi = vo;
            vo = tv;
        }
        // perm: sorted indices of binned visibilities first (sentinels sort last)
    }
    MG_TRY(exclusive_scan(hist, g.nkeys, offsets, partial, true, s));
    return MIGRID_OK;
}

}  // namespace migrid

using namespace migrid;

extern "C" int64_t migrid_bins_per_axis(int64_t img_dim, int64_t gcf_dim, int32_t bin)
{
    if (bin <= 0) return 0;
    return (img_dim + gcf_dim - 2) / bin + 1;
}

extern "C" int migrid_ctx_bin_sort(migrid_ctx* ctx, int32_t in_type, const void* uv, int64_t npts,
                                   int64_t img_dim, int64_t gcf_dim, int32_t bin, int32_t* perm,
                                   int32_t* offsets, int64_t* count, void* stream)
{
    clear_error();
    if (!ctx) return fail(MIGRID_EINVAL, "ctx is NULL");
    MG_TRY(validate_shape(npts, img_dim, gcf_dim));
    if (bin < 1) return fail(MIGRID_EINVAL, "bin=%d must be >= 1", bin);
    if (in_type != MIGRID_F64 && in_type != MIGRID_F32)
        return fail(MIGRID_EINVAL, "in_type=%d", in_type);
    if (!perm || !offsets || (npts > 0 && !uv)) return fail(MIGRID_EINVAL, "NULL buffer");
    const BinGeom g = bin_geom(img_dim, gcf_dim, bin);
    if (g.nkeys >= (int64_t)1 << 31) return fail(MIGRID_EINVAL, "too many bins (%lld)", (long long)g.nkeys);
    MG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    MG_TRY(ctx->bins.ensure(bin_scratch_bytes(npts, g)));
    MG_TRY(bin_sort(g, in_type, uv, npts, ctx->bins.p, perm, offsets, s));
    int32_t total = 0;
    MG_HIP(hipMemcpyAsync(&total, offsets + g.nkeys, 4, hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
    if (count) *count = total;
    return MIGRID_OK;
}
