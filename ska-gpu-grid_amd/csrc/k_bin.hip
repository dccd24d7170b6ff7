// Bin sort of visibilities by (job, sub-kernel k, bin_y, bin_x).
//
// Replaces the reference's host qsort (grid.cu:467-475), vis2ints
// (grid_gpu.cu:352-365) and set_bookmarks (:366-390) with an on-device,
// stable sort over the visibilities of any number of jobs at once (their
// elements concatenated, each job owning a range of keys):
//   1. k_keys     per visibility: q = floor(uv) + S/2, k = 8*sub_y + sub_x,
//                 key = job base + (k*nb + q_y/H)*nb + q_x/H, or the job's
//                 sentinel (base + nkeys) for a visibility that touches no
//                 pixel; one block row per job (blockIdx.y);
//   2. LSD radix sort of (key, element) in 8-bit digits; each pass ranks a
//      block's 2048 elements stably with wave64 ballots (a lane's peers are
//      the lanes whose digit matches in all 8 bit-ballots; its rank is the
//      popcount of the peers below it, __builtin_amdgcn_mbcnt) and an LDS
//      prefix over the block's waves; the first pass takes the element index
//      as its value;
//   3. CSR offsets (the bookmarks) from the sorted keys, offsets[k] = number
//      of elements with key < k, written once per key: the key space is cut
//      into chunks of 2048 keys, each chunk's first sorted position is found
//      from the element where it begins, and per chunk the first position of
//      every present key is marked in LDS and a suffix minimum fills the
//      keys that are absent -- no histogram of atomics.
// The permutation is the stable order, so it is bit-identical to a serial
// counting sort of the same keys.
#include <hip/hip_runtime.h>

#include <climits>

#include "bin.hpp"
#include "device_common.hpp"
#include "internal.hpp"
#include "table.hpp"

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
__global__ __launch_bounds__(256) void k_keys(const BinJob* __restrict__ jobs,
                                              uint32_t* __restrict__ keys)
{
    const BinJob jb = jobs[blockIdx.y];
    const TI* __restrict__ uv = (const TI*)jb.uv;
    const int nb = jb.nb;
    const uint32_t nkeys = (uint32_t)kSubKernels * nb * nb;
    const uint32_t kbase = (uint32_t)jb.k_base;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < jb.npts;
         n += (int64_t)gridDim.x * blockDim.x)
        keys[jb.e_base + n] = kbase + bin_key<TI>(uv[2 * n], uv[2 * n + 1], jb.N, jb.S, jb.H, nb, nkeys);
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
// each round is ranked with ballots plus an LDS prefix over the 4 waves.  The
// ranked elements are placed in LDS in digit order first (the block's digit
// counts, from k_rs_hist, give each digit's local start), then written out
// by consecutive threads, so each digit's run of the block goes to global
// memory as contiguous stores rather than one 4-byte store per element.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift, const int32_t* __restrict__ base_offs,
    const int32_t* __restrict__ counts)
{
    __shared__ int wave_cnt[4][256];
    __shared__ int running[256];
    __shared__ int blk_base[256];
    __shared__ int loc_start[256];
    __shared__ int wave_sum[4];
    __shared__ uint32_t sk[kRsTile];
    __shared__ uint32_t sv[kRsTile];
    const int t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    running[t] = 0;
    blk_base[t] = base_offs[(int64_t)t * gridDim.x + blockIdx.x];
    {
        // exclusive prefix of the block's digit counts (thread t = digit t)
        const int c = counts[(int64_t)t * gridDim.x + blockIdx.x];
        int x = c;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wave_sum[w] = x;
        __syncthreads();
        int off = 0;
        for (int ww = 0; ww < w; ww++) off += wave_sum[ww];
        loc_start[t] = off + x - c;
    }
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
            val = vin ? vin[i] : (uint32_t)i;  // the first pass: the element index
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
            const int lp = loc_start[d] + wave_cnt[w][d] + rank;  // block-local sorted position
            sk[lp] = key;
            sv[lp] = val;
        }
        __syncthreads();
    }
    const int64_t rem = n - base;
    const int cnt_blk = rem < kRsTile ? (int)rem : kRsTile;
    for (int e = t; e < cnt_blk; e += kRsThreads) {
        const uint32_t key = sk[e];
        const int d = (key >> shift) & 255;
        const int64_t pos = (int64_t)blk_base[d] + (e - loc_start[d]);
        kout[pos] = key;
        vout[pos] = sv[e];
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

// Inclusive suffix minimum over a wave's lanes (lane l: min over lanes l..63).
__device__ __forceinline__ int wave_suffix_min(int x)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_down(x, o, 64);
        if (lane + o < 64) x = min(x, y);
    }
    return x;
}

// ---- CSR offsets: offsets[j] = lower bound of key j in the sorted keys ------
// Key space in chunks of kScanTile keys (one workgroup each).  First, for
// every chunk c, start[c] = the first sorted position whose key is >= c C
// (start is filled with the element count beforehand): the element at p
// writes the chunks whose first key lies in (sk[p - 1], sk[p]].
__global__ __launch_bounds__(256) void k_chunk_starts(const uint32_t* __restrict__ sk, int64_t n,
                                                      int32_t* __restrict__ start)
{
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t hi = sk[p] / kScanTile;
        const int64_t lo = p == 0 ? 0 : (int64_t)(sk[p - 1] / kScanTile) + 1;
        for (int64_t c = lo; c <= hi; c++) start[c] = (int32_t)p;
    }
}

// Then each workgroup writes its chunk's offsets once: the first position of
// every key present in the chunk goes to LDS (positions [start[c], start[c +
// 1]) hold exactly the chunk's keys), and a suffix minimum over the chunk,
// seeded with start[c + 1], gives every key's lower bound.  A chunk holding
// more than kLbScan elements (many visibilities on a few keys) takes a binary
// search per key instead of the scan.
constexpr int64_t kLbScan = 16 * kScanTile;

__global__ __launch_bounds__(256) void k_lower_bounds(const uint32_t* __restrict__ sk,
                                                      int64_t K, const int32_t* __restrict__ start,
                                                      int32_t* __restrict__ offs)
{
    static_assert(kScanItems == 8, "two 16-byte stores per thread");
    __shared__ int32_t s_first[kScanTile];
    __shared__ int ws[4];
    const int tid = threadIdx.x, lane = tid & 63;
    const int64_t c = blockIdx.x;
    const int64_t j0 = c * kScanTile;
    const int32_t i0 = start[c], i1 = start[c + 1];
    if ((int64_t)i1 - i0 > kLbScan) {
        for (int e = tid; e < kScanTile && j0 + e < K; e += 256) {
            const uint32_t j = (uint32_t)(j0 + e);
            int32_t lo = i0, hi = i1;
            while (lo < hi) {
                const int32_t mid = lo + (hi - lo) / 2;
                if (sk[mid] < j)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            offs[j0 + e] = lo;
        }
        return;
    }
    for (int e = tid; e < kScanTile; e += 256) s_first[e] = i1;
    __syncthreads();
    for (int32_t p = i0 + tid; p < i1; p += 256) {
        const uint32_t k = sk[p];
        if (p == 0 || sk[p - 1] != k) s_first[k - (uint32_t)j0] = p;
    }
    __syncthreads();
    // thread t: keys [8 t, 8 t + 8) of the chunk
    int v[kScanItems];
    int m = INT_MAX;
    for (int e = 0; e < kScanItems; e++) {
        v[e] = s_first[kScanItems * tid + e];
        m = min(m, v[e]);
    }
    const int inc = wave_suffix_min(m);
    if (lane == 0) ws[tid >> 6] = inc;
    __syncthreads();
    const int later_in_wave = __shfl_down(inc, 1, 64);
    int run = lane < 63 ? later_in_wave : INT_MAX;
    for (int w = (tid >> 6) + 1; w < 4; w++) run = min(run, ws[w]);
    for (int e = kScanItems - 1; e >= 0; e--) {
        run = min(run, v[e]);
        v[e] = run;
    }
    const int64_t base = j0 + kScanItems * tid;
    if (((uintptr_t)offs & 15) == 0 && base + kScanItems <= K) {
        *(int4*)(offs + base) = make_int4(v[0], v[1], v[2], v[3]);
        *(int4*)(offs + base + 4) = make_int4(v[4], v[5], v[6], v[7]);
    } else {
        for (int e = 0; e < kScanItems; e++)
            if (base + e < K) offs[base + e] = v[e];
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

namespace {
size_t align64(size_t b) { return (b + 63) & ~(size_t)63; }
}  // namespace

int put_binjobs(BinJob* dst, const BinJob* src, int n, hipStream_t s)
{
    return put_table(dst, src, n, s);
}

size_t bin_scratch_bytes(int64_t E, int64_t K)
{
    const int64_t nblk = (E + kRsTile - 1) / kRsTile;
    size_t b = 4 * align64((size_t)(E > 0 ? E : 1) * 4);  // keys/values double buffers
    b += 2 * align64((size_t)256 * (nblk + 1) * 4);        // digit counts + offsets
    const int64_t big = K > 256 * (nblk + 1) ? K : 256 * (nblk + 1);
    b += align64((size_t)scan_scratch_elems(big) * 4);     // scan partials
    return b;
}

int bin_sort(const BinJob* jobs, int njobs, int64_t max_npts, int64_t E, int64_t K, int in_type,
             void* scratch, int32_t* perm, int32_t* offsets, hipStream_t s)
{
    if (E > 0) {
        uint32_t* k0 = sort_key_buffer(scratch);
        const int64_t per_job = (max_npts + 255) / 256;
        const int64_t cap = njobs < 16384 ? 16384 / njobs : 1;
        const int bx = (int)(per_job < 1 ? 1 : per_job < cap ? per_job : cap);
        if (in_type == MIGRID_F64)
            hipLaunchKernelGGL(k_keys<double>, dim3(bx, njobs), dim3(256), 0, s, jobs, k0);
        else
            hipLaunchKernelGGL(k_keys<float>, dim3(bx, njobs), dim3(256), 0, s, jobs, k0);
        MG_HIP(hipGetLastError());
    }
    return sort_keys(E, K, scratch, perm, offsets, s);
}

uint32_t* sort_key_buffer(void* scratch) { return (uint32_t*)scratch; }

int sort_keys(int64_t E, int64_t K, void* scratch, int32_t* perm, int32_t* offsets, hipStream_t s)
{
    const int64_t nblk = (E + kRsTile - 1) / kRsTile;
    char* p = (char*)scratch;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += align64(bytes);
        return r;
    };
    const size_t eb = (size_t)(E > 0 ? E : 1) * 4;
    uint32_t* k0 = (uint32_t*)take(eb);
    uint32_t* v0 = (uint32_t*)take(eb);
    uint32_t* k1 = (uint32_t*)take(eb);
    uint32_t* v1 = (uint32_t*)take(eb);
    int32_t* counts = (int32_t*)take((size_t)256 * (nblk + 1) * 4);
    int32_t* coffs = (int32_t*)take((size_t)256 * (nblk + 1) * 4);
    const int64_t big = K > 256 * (nblk + 1) ? K : 256 * (nblk + 1);
    int32_t* partial = (int32_t*)take((size_t)scan_scratch_elems(big) * 4);

    if (E == 0) return fill_u32(offsets, 0u, K, s);
    {
        int bits = 1;
        while (((int64_t)1 << bits) < K) bits++;  // keys are < K
        const int passes = (bits + 7) / 8;
        uint32_t *ki = k0, *vi = nullptr, *ko = k1, *vo = v1;
        uint32_t* vspare = v0;
        for (int pass = 0; pass < passes; pass++) {
            const int shift = 8 * pass;
            if (pass == passes - 1) vo = (uint32_t*)perm;  // the last pass writes the permutation
            hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(kRsThreads), 0, s, ki, E, shift, counts);
            MG_TRY(exclusive_scan(counts, 256 * nblk, coffs, partial, false, s));
            hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(kRsThreads), 0, s, ki, vi, ko, vo, E,
                               shift, coffs, counts);
            MG_HIP(hipGetLastError());
            uint32_t* tk = ki;
            ki = ko;
            ko = tk;
            uint32_t* next_vo = vi ? vi : vspare;
            vi = vo;
            vo = next_vo;
        }
        // ki: the sorted keys; partial: the chunks' first positions
        const int64_t nchunk = (K + kScanTile - 1) / kScanTile;
        MG_TRY(fill_u32((uint32_t*)partial, (uint32_t)E, nchunk + 1, s));
        const int mb = (int)((E + 255) / 256 < 8192 ? (E + 255) / 256 : 8192);
        hipLaunchKernelGGL(k_chunk_starts, dim3(mb), dim3(256), 0, s, ki, E, partial);
        hipLaunchKernelGGL(k_lower_bounds, dim3(nchunk), dim3(256), 0, s, ki, K, partial, offsets);
        MG_HIP(hipGetLastError());
    }
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
    if (g.nkeys >= (int64_t)1 << 31 || npts >= (int64_t)1 << 31)
        return fail(MIGRID_EINVAL, "too many bins (%lld) or visibilities", (long long)g.nkeys);
    MG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const int64_t K = g.nkeys + 1;  // the job's keys and its sentinel
    MG_TRY(ctx->bins.ensure(bin_scratch_bytes(npts, K) + 256));
    BinJob jb;
    jb.uv = uv;
    jb.npts = npts;
    jb.e_base = 0;
    jb.k_base = 0;
    jb.N = g.N;
    jb.S = g.S;
    jb.H = g.H;
    jb.nb = g.nb;
    BinJob* djob = (BinJob*)ctx->bins.p;
    MG_TRY(put_binjobs(djob, &jb, 1, s));
    MG_TRY(bin_sort(djob, 1, npts, npts, K, in_type, (char*)ctx->bins.p + 256, perm, offsets, s));
    int32_t total = 0;
    MG_HIP(hipMemcpyAsync(&total, offsets + g.nkeys, 4, hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
    if (count) *count = total;
    return MIGRID_OK;
}
