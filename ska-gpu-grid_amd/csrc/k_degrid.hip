// Degridding: the adjoint of gridding (grid -> visibilities).  No counterpart
// in the reference (the step after its path, for prediction / major cycles).
//
//   vals[n] = sum over the in-grid taps of visibility n of
//             conj(G[k][r][c]) * grid[main_y - S/2 + 1 + r][main_x - S/2 + 1 + c]
//
// with the gridder's index math, so <grid(v), g> = <v, degrid(g)>.
//
// MI355X layout: a visibility belongs to the tile that holds its window's
// corner, q = floor(uv) + S/2 binned by the tile edge T (the gridder's q),
// and the visibilities are sorted by (tile, sub-kernel).  A tile's run is cut
// into work items of at most `target` visibilities, so a dense tile is shared
// by many workgroups.  Persistent workgroups take items in turn: stage the
// item's tile and its (S - 1)-pixel halo of the grid in LDS (zeros off the
// grid), then each wave takes groups of V consecutive visibilities: its lanes
// walk the sub-kernel's columns, reading a GCF row (coalesced) and the
// matching staged grid row of each visibility, and a 64-lane reduction gives
// each value.  A group whose visibilities share the sub-kernel (the common
// case in a dense tile, being sorted by it) reads each GCF row once for all V.
// Every visibility is computed by exactly one wave, so no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bin.hpp"
#include "device_common.hpp"
#include "internal.hpp"

namespace migrid {
namespace {

// The automatic tile edge: the largest of 64, 32, 16 whose staged tile (with
// its S - 1 halo, gsz bytes a pixel) fits a CU's 160 KB of LDS; 0 if none
// fits.  Larger tiles mean fewer stagings and longer runs of one sub-kernel,
// which outweigh a second workgroup per CU (measured at S = 64, FP32).
constexpr int auto_tile(int S, int gsz)
{
    for (int t = 64; t >= 16; t /= 2)
        if ((t + S - 1) * (t + S - 1) * gsz <= 160 * 1024) return t;
    return 0;
}

// Threads of a degridding workgroup: 16 waves (the staging loop strides by it).
constexpr int kDegridThreads = 1024;

struct DegridItem {
    int32_t tile, lo, hi, pad;  // sorted positions [lo, hi) of one tile
};

struct DegridArgs {
    const void* uv;
    const void* gtab;  // 64 * S^2 complex in the product type
    const void* grid;
    void* out;
    const int4* rec;  // per sorted position: {n, tile offset of the window, k, 0}
    const DegridItem* items;
    const int32_t* nitems;  // device count of items
    int N, S, T, ntx;
    int planar;
};

template <typename T>
struct C2 {
    T re, im;
};

// Sum over the 64 lanes of a wave, left in lane 63, by DPP moves (no LDS):
// pairs, quads (quad_perm), halves of a row (row_half_mirror), rows
// (row_mirror), then row 0 into row 1 and row 2 into row 3 (row_bcast:15),
// then rows 0-1 into rows 2-3 (row_bcast:31).  Every lane must be active.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWS, 0xf, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f(double x)
{
    const int64_t b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}
template <typename T>
__device__ __forceinline__ T wave_sum63(T x)
{
    x += dpp_f<0xb1, 0xf>(x);   // quad_perm [1,0,3,2]
    x += dpp_f<0x4e, 0xf>(x);   // quad_perm [2,3,0,1]
    x += dpp_f<0x141, 0xf>(x);  // row_half_mirror
    x += dpp_f<0x140, 0xf>(x);  // row_mirror
    x += dpp_f<0x142, 0xa>(x);  // row_bcast:15 into rows 1, 3
    x += dpp_f<0x143, 0xc>(x);  // row_bcast:31 into rows 2, 3
    return x;
}

// Sort keys: (tile of q, k), or the sentinel ntiles*64 for a window that
// misses the grid.
template <typename TI>
__global__ __launch_bounds__(256) void k_degrid_keys(const TI* __restrict__ uv, int64_t npts, int N,
                                                     int S, int T, int ntx,
                                                     uint32_t* __restrict__ keys)
{
    const uint32_t sentinel = (uint32_t)ntx * ntx * kSubKernels;
    const int qmax = N + S - 2;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < npts;
         n += (int64_t)gridDim.x * blockDim.x) {
        const TI u = uv[2 * n], v = uv[2 * n + 1];
        const TI fu = floor_t(u), fv = floor_t(v);
        const int sx = (int)floor_t((TI)kOversample * (u - fu));
        const int sy = (int)floor_t((TI)kOversample * (v - fv));
        const TI qxf = fu + (TI)(S / 2), qyf = fv + (TI)(S / 2);
        uint32_t key = sentinel;
        if (qxf >= 0 && qyf >= 0 && qxf <= (TI)qmax && qyf <= (TI)qmax) {
            const int qx = (int)qxf, qy = (int)qyf;
            key = ((uint32_t)(qy / T) * ntx + qx / T) * kSubKernels + (kOversample * sy + sx);
        }
        keys[n] = key;
    }
}

// Items per tile: ceil(count / target).
__global__ __launch_bounds__(256) void k_degrid_item_counts(const int32_t* __restrict__ offs,
                                                            int64_t ntiles, int target,
                                                            int32_t* __restrict__ ni)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntiles;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t cnt = offs[(t + 1) * kSubKernels] - offs[t * kSubKernels];
        ni[t] = (cnt + target - 1) / target;
    }
}

// A tile's run split into ni[t] nearly equal items starting at istart[t].
__global__ __launch_bounds__(256) void k_degrid_items(const int32_t* __restrict__ offs,
                                                      const int32_t* __restrict__ ni,
                                                      const int32_t* __restrict__ istart,
                                                      int64_t ntiles, DegridItem* __restrict__ items)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntiles;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t m = ni[t];
        const int32_t a = offs[t * kSubKernels];
        const int64_t cnt = offs[(t + 1) * kSubKernels] - a;
        DegridItem* dst = items + istart[t];
        for (int32_t i = 0; i < m; ++i) {
            DegridItem it;
            it.tile = (int32_t)t;
            it.lo = a + (int32_t)(cnt * i / m);
            it.hi = a + (int32_t)(cnt * (i + 1) / m);
            it.pad = 0;
            dst[i] = it;
        }
    }
}

// Records in sorted order for the visibilities that fall in a tile (the
// first offs_end[0] sorted positions): the index n, the offset of the
// window's corner in the staged tile (W pixels a row) and the sub-kernel k.
template <typename TI>
__global__ __launch_bounds__(256) void k_degrid_pack(const int32_t* __restrict__ perm,
                                                     const TI* __restrict__ uv,
                                                     const int32_t* __restrict__ offs_end, int S,
                                                     int T, int4* __restrict__ rec)
{
    const int64_t e = *offs_end;
    const int W = T + S - 1;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < e;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int n = perm[j];
        const TI u = uv[2 * (int64_t)n], v = uv[2 * (int64_t)n + 1];
        const TI fu = floor_t(u), fv = floor_t(v);
        const int k = kOversample * (int)floor_t((TI)kOversample * (v - fv)) +
                      (int)floor_t((TI)kOversample * (u - fu));
        const int qx = (int)fu + S / 2, qy = (int)fv + S / 2;  // in [0, N + S - 2]
        const int xr = qx - (qx / T) * T, yr = qy - (qy / T) * T;
        rec[j] = make_int4(n, yr * W + xr, k, 0);
    }
}

// Sums of conj(G[k]) x window over the sub-kernel rows of this lane for VV
// visibilities sharing gk (this lane's column of sub-kernel k) at staged
// offsets off[] (this lane's column included).  Rows go in blocks of U; the
// next block's coefficients are loaded while the current block computes.
template <int VV, int U, typename G, typename GS, typename TA>
__device__ __forceinline__ void degrid_rows(const G* __restrict__ gk, const GS* tile,
                                            const int* off, int rr, int R, int S, int W, TA* ar,
                                            TA* ai)
{
    const int step = U * R;
    if (S % step == 0) {  // every active lane has S / R rows, in whole blocks
        const int nb = S / step;  // the same for every lane
        const G* __restrict__ gp = gk + rr * S;
        const GS* tp = tile + rr * W;
        G gn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) gn[u] = gp[u * R * S];
        for (int b = 0; b < nb; ++b) {
            G g[U];
#pragma unroll
            for (int u = 0; u < U; ++u) g[u] = gn[u];
            if (b + 1 < nb)
#pragma unroll
                for (int u = 0; u < U; ++u) gn[u] = gp[((b + 1) * step + u * R) * S];
            GS x[U][VV];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int i = 0; i < VV; ++i) x[u][i] = tp[off[i] + (b * step + u * R) * W];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const TA gr = (TA)g[u].x, gi = (TA)g[u].y;
#pragma unroll
                for (int i = 0; i < VV; ++i) {
                    const TA xr = (TA)x[u][i].re, xi = (TA)x[u][i].im;
                    ar[i] = fma(gr, xr, ar[i]);
                    ar[i] = fma(gi, xi, ar[i]);
                    ai[i] = fma(gr, xi, ai[i]);
                    ai[i] = fma(-gi, xr, ai[i]);
                }
            }
        }
    } else {
        for (int r = rr; r < S; r += R) {
            const G g = gk[r * S];
            const TA gr = (TA)g.x, gi = (TA)g.y;
#pragma unroll
            for (int i = 0; i < VV; ++i) {
                const GS x = tile[off[i] + r * W];
                const TA xr = (TA)x.re, xi = (TA)x.im;
                ar[i] = fma(gr, xr, ar[i]);
                ar[i] = fma(gi, xi, ar[i]);
                ai[i] = fma(gr, xi, ai[i]);
                ai[i] = fma(-gi, xr, ai[i]);
            }
        }
    }
}

// G: coefficient pair type, GS: staged grid pair type, TA: accumulation,
// V: visibilities per wave pass; SC, WC: the support and the staged tile's
// edge as compile-time constants (0: read from the arguments), which turns
// the row loop's addressing into immediate offsets.
template <int V, int SC, int WC, typename G, typename GS, typename TA, typename TI, typename TO>
__global__ __launch_bounds__(kDegridThreads) void k_degrid(DegridArgs a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    GS* tile = (GS*)smem_raw;
    const int N = a.N, S = SC ? SC : a.S, T = a.T;
    const int W = WC ? WC : T + S - 1;
    const TO* __restrict__ grid = (const TO*)a.grid;
    const int64_t NN = (int64_t)N * N;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int nw = kDegridThreads / 64;
    const int R = 64 / S;  // rows a pass covers (S <= 64)
    const int c = lane % S, rr = lane / S;
    const bool active = rr < R;
    const G* __restrict__ gtab = (const G*)a.gtab;
    TI* __restrict__ out = (TI*)a.out;
    const int total = *a.nitems;
    constexpr int U = sizeof(TA) == 8 ? 2 : 4;  // rows per block
    // staged pixels a thread loads per pass: all of them for a known edge,
    // at most 5 double pairs (registers)
    constexpr int kAll = WC ? (WC * WC + kDegridThreads - 1) / kDegridThreads : 4;
    constexpr int kStagePer = sizeof(GS) == 16 && kAll > 5 ? 5 : kAll;
    for (int it = blockIdx.x; it < total; it += gridDim.x) {
        const DegridItem item = a.items[it];
        const int tx = item.tile % a.ntx, ty = item.tile / a.ntx;
        const int X0 = tx * T - (S - 1), Y0 = ty * T - (S - 1);
        // all of a thread's loads in flight before its stores (one memory
        // latency per pass rather than one per pixel)
        for (int e0 = threadIdx.x; e0 < W * W; e0 += kStagePer * kDegridThreads) {
            GS g[kStagePer];
#pragma unroll
            for (int p = 0; p < kStagePer; ++p) {
                const int e = e0 + p * kDegridThreads;
                const int yy = e / W, xx = e - yy * W;
                const int y = Y0 + yy, x = X0 + xx;
                g[p].re = 0;
                g[p].im = 0;
                if (e < W * W && x >= 0 && y >= 0 && x < N && y < N) {
                    const int64_t q = (int64_t)y * N + x;
                    g[p].re = (decltype(g[p].re))(a.planar ? grid[q] : grid[2 * q]);
                    g[p].im = (decltype(g[p].im))(a.planar ? grid[NN + q] : grid[2 * q + 1]);
                }
            }
#pragma unroll
            for (int p = 0; p < kStagePer; ++p)
                if (e0 + p * kDegridThreads < W * W) tile[e0 + p * kDegridThreads] = g[p];
        }
        __syncthreads();
        // this wave's contiguous share of the item, in groups of V
        const int len = item.hi - item.lo;
        const int w0 = item.lo + (int)((int64_t)len * wave / nw);
        const int w1 = item.lo + (int)((int64_t)len * (wave + 1) / nw);
        int4 recn = make_int4(0, 0, 0, 0);
        if (w0 < w1) recn = a.rec[min(w0 + lane % V, w1 - 1)];
        for (int j0 = w0; j0 < w1; j0 += V) {
            const int4 rc = recn;  // lane i: visibility j0 + min(i, nv - 1)
            const int nv = w1 - j0 < V ? w1 - j0 : V;
            if (j0 + V < w1) recn = a.rec[min(j0 + V + lane % V, w1 - 1)];
            int n[V], off[V], k[V];
#pragma unroll
            for (int i = 0; i < V; ++i) {
                n[i] = __builtin_amdgcn_readlane(rc.x, i);
                off[i] = __builtin_amdgcn_readlane(rc.y, i) + c;
                k[i] = __builtin_amdgcn_readlane(rc.z, i);
            }
            TA ar[V], ai[V];
#pragma unroll
            for (int i = 0; i < V; ++i) ar[i] = ai[i] = 0;
            if (active) {
                if (k[0] == k[V - 1]) {  // sorted by k: the whole group shares it
                    degrid_rows<V, U>(gtab + (int64_t)k[0] * S * S + c, tile, off, rr, R, S, W,
                                      ar, ai);
                } else {
#pragma unroll
                    for (int i = 0; i < V; ++i)
                        degrid_rows<1, U>(gtab + (int64_t)k[i] * S * S + c, tile, off + i, rr, R,
                                          S, W, ar + i, ai + i);
                }
            }
#pragma unroll
            for (int i = 0; i < V; ++i) {
                ar[i] = wave_sum63(ar[i]);
                ai[i] = wave_sum63(ai[i]);
            }
            if (lane == 63)
#pragma unroll
                for (int i = 0; i < V; ++i)
                    if (i < nv) {
                        out[2 * (int64_t)n[i]] = (TI)ar[i];
                        out[2 * (int64_t)n[i] + 1] = (TI)ai[i];
                    }
        }
        __syncthreads();  // the tile is restaged by the next item
    }
}

template <typename TI, typename G>
__global__ __launch_bounds__(256) void k_degrid_convert(const TI* __restrict__ src,
                                                        G* __restrict__ dst, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        G g;
        g.x = (decltype(g.x))src[2 * i];
        g.y = (decltype(g.y))src[2 * i + 1];
        dst[i] = g;
    }
}

int blocks_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : b > 8192 ? 8192 : b);
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

template <int V, int SC, int WC, typename G, typename GS, typename TA, typename TI, typename TO>
int run_degrid_v(const DegridArgs& a, size_t lds, int wg_per_cu, hipStream_t s)
{
    auto kern = k_degrid<V, SC, WC, G, GS, TA, TI, TO>;
    MG_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    int per_cu = 0;
    MG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kDegridThreads, lds));
    if (per_cu < 1) return fail(MIGRID_ENOTSUP, "degridding kernel does not fit (lds=%zu)", lds);
    if (wg_per_cu > 0 && wg_per_cu < per_cu) per_cu = wg_per_cu;
    int dev = 0, ncu = 0;
    MG_HIP(hipGetDevice(&dev));
    MG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    hipLaunchKernelGGL(kern, dim3((unsigned)(ncu * per_cu)), dim3(kDegridThreads), lds, s, a);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

// The support-specialised kernels (S = 8, 16, 32, 64 with the automatic
// tile edge) or the general one.
template <int V, typename G, typename GS, typename TA, typename TI, typename TO>
int run_degrid_s(const DegridArgs& a, bool special, size_t lds, int wpc, hipStream_t s)
{
    constexpr int g = (int)sizeof(GS);
    if (special) {
        switch (a.S) {
        case 8: return run_degrid_v<V, 8, auto_tile(8, g) + 7, G, GS, TA, TI, TO>(a, lds, wpc, s);
        case 16: return run_degrid_v<V, 16, auto_tile(16, g) + 15, G, GS, TA, TI, TO>(a, lds, wpc, s);
        case 32: return run_degrid_v<V, 32, auto_tile(32, g) + 31, G, GS, TA, TI, TO>(a, lds, wpc, s);
        case 64: return run_degrid_v<V, 64, auto_tile(64, g) + 63, G, GS, TA, TI, TO>(a, lds, wpc, s);
        default: break;
        }
    }
    return run_degrid_v<V, 0, 0, G, GS, TA, TI, TO>(a, lds, wpc, s);
}

template <typename G, typename GS, typename TA, typename TI, typename TO>
int run_degrid(const DegridArgs& a, int V, bool special, size_t lds, int wpc, hipStream_t s)
{
    switch (V) {
    case 1: return run_degrid_s<1, G, GS, TA, TI, TO>(a, special, lds, wpc, s);
    case 2: return run_degrid_s<2, G, GS, TA, TI, TO>(a, special, lds, wpc, s);
    default: return run_degrid_s<4, G, GS, TA, TI, TO>(a, special, lds, wpc, s);
    }
}

}  // namespace
}  // namespace migrid

using namespace migrid;

extern "C" int migrid_ctx_degrid(migrid_ctx* ctx, const migrid_job* job, void* vals_out,
                                 const migrid_opts* opts_in, void* stream)
{
    clear_error();
    if (!ctx || !job) return fail(MIGRID_EINVAL, "NULL argument");
    migrid_opts o;
    if (opts_in)
        o = *opts_in;
    else
        migrid_opts_default(&o);
    MG_TRY(validate_shape(job->npts, job->img_dim, job->gcf_dim));
    if (job->gcf_dim > 64) return fail(MIGRID_ENOTSUP, "degridding supports S <= 64");
    if (job->img_dim + job->gcf_dim - 2 >= 65536) return fail(MIGRID_ENOTSUP, "N + S - 2 >= 65536");
    if (o.gcf_mode != MIGRID_GCF_TABLE) return fail(MIGRID_ENOTSUP, "degridding reads a GCF table");
    if (o.arith == MIGRID_ARITH_F32_BF16GCF) return fail(MIGRID_ENOTSUP, "degridding: no BF16 GCF");
    const bool in64 = job->in_type == MIGRID_F64, out64 = job->out_type == MIGRID_F64;
    if ((job->in_type != MIGRID_F64 && job->in_type != MIGRID_F32) ||
        (job->out_type != MIGRID_F64 && job->out_type != MIGRID_F32))
        return fail(MIGRID_EINVAL, "types %d/%d", job->in_type, job->out_type);
    if (in64 && !out64 && o.arith == MIGRID_ARITH_NATIVE)
        return fail(MIGRID_ENOTSUP, "double visibilities from a float grid need FP32 arithmetic");
    if (!job->grid || !job->gcf || (job->npts > 0 && (!job->uv || !vals_out)))
        return fail(MIGRID_EINVAL, "NULL buffer");
    const int64_t npts = job->npts;
    const int N = (int)job->img_dim, S = (int)job->gcf_dim;
    const bool f32 = o.arith == MIGRID_ARITH_F32 || !in64;  // products in float
    // sums in double for double visibilities with native arithmetic, and for
    // float visibilities from a double grid (the <double2, float2> pairing);
    // the staged grid is kept in the type of the sums
    const bool acc64 = (!f32) || (!in64 && out64);
    const size_t gsz = acc64 ? 16 : 8;
    auto tile_lds = [&](int t) { return (size_t)(t + S - 1) * (t + S - 1) * gsz; };
    int T = auto_tile(S, (int)gsz);
    if (o.tile > 0) {
        if (o.tile != 16 && o.tile != 32 && o.tile != 64)
            return fail(MIGRID_EINVAL, "degridding tile %d (16, 32 or 64)", o.tile);
        T = o.tile;
    }
    // the specialised kernels assume the automatic tile edge
    const bool special = o.tile <= 0 && (S == 8 || S == 16 || S == 32 || S == 64);
    if (!T || tile_lds(T) > 160 * 1024)
        return fail(MIGRID_ENOTSUP, "degridding tile does not fit in LDS");
    const size_t lds = tile_lds(T);
    const int V = o.vis_per_wave > 0 ? o.vis_per_wave : 4;
    if (V != 1 && V != 2 && V != 4)
        return fail(MIGRID_EINVAL, "degridding: vis_per_wave %d (1, 2 or 4)", V);
    const int target = 2048;  // visibilities per work item
    const int ntx = (N + S - 2) / T + 1;
    const int64_t ntiles = (int64_t)ntx * ntx;
    const int64_t K = ntiles * kSubKernels + 1;
    if (K >= ((int64_t)1 << 31) || npts >= ((int64_t)1 << 31))
        return fail(MIGRID_ENOTSUP, "too many tiles or visibilities");
    MG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t ce_in = in64 ? 16 : 8;
    MG_TRY(launch_zero(vals_out, (size_t)npts * ce_in, s));  // windows that miss the grid
    if (npts == 0) return MIGRID_OK;
    // arena: permutation, CSR offsets, items per tile and their starts, scan
    // partials, the items, the records, the GCF in the product type
    const bool convert = f32 && in64;
    const int64_t max_items = npts / target + ntiles;
    const size_t b_perm = up256((size_t)npts * 4), b_offs = up256((size_t)K * 4);
    const size_t b_ni = up256((size_t)ntiles * 4), b_is = up256((size_t)(ntiles + 1) * 4);
    const size_t b_part = up256((size_t)scan_scratch_elems(ntiles) * 4);
    const size_t b_items = up256((size_t)max_items * sizeof(DegridItem));
    const size_t b_rec = up256((size_t)npts * sizeof(int4));
    const size_t gtab_bytes = convert ? (size_t)kSubKernels * S * S * 8 : 0;
    MG_TRY(ctx->records.ensure(b_perm + b_offs + b_ni + b_is + b_part + b_items + b_rec +
                               gtab_bytes));
    char* ar = (char*)ctx->records.p;
    int32_t* perm = (int32_t*)ar;
    ar += b_perm;
    int32_t* offs = (int32_t*)ar;
    ar += b_offs;
    int32_t* ni = (int32_t*)ar;
    ar += b_ni;
    int32_t* istart = (int32_t*)ar;
    ar += b_is;
    int32_t* partial = (int32_t*)ar;
    ar += b_part;
    DegridItem* items = (DegridItem*)ar;
    ar += b_items;
    int4* rec = (int4*)ar;
    ar += b_rec;
    void* gtab = ar;
    MG_TRY(ctx->bins.ensure(bin_scratch_bytes(npts, K)));
    if (in64)
        hipLaunchKernelGGL(k_degrid_keys<double>, dim3(blocks_for(npts)), dim3(256), 0, s,
                           (const double*)job->uv, npts, N, S, T, ntx, sort_key_buffer(ctx->bins.p));
    else
        hipLaunchKernelGGL(k_degrid_keys<float>, dim3(blocks_for(npts)), dim3(256), 0, s,
                           (const float*)job->uv, npts, N, S, T, ntx, sort_key_buffer(ctx->bins.p));
    MG_HIP(hipGetLastError());
    MG_TRY(sort_keys(npts, K, ctx->bins.p, perm, offs, s));
    if (in64)
        hipLaunchKernelGGL(k_degrid_pack<double>, dim3(blocks_for(npts)), dim3(256), 0, s, perm,
                           (const double*)job->uv, offs + ntiles * kSubKernels, S, T, rec);
    else
        hipLaunchKernelGGL(k_degrid_pack<float>, dim3(blocks_for(npts)), dim3(256), 0, s, perm,
                           (const float*)job->uv, offs + ntiles * kSubKernels, S, T, rec);
    MG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_degrid_item_counts, dim3(blocks_for(ntiles)), dim3(256), 0, s, offs,
                       ntiles, target, ni);
    MG_HIP(hipGetLastError());
    MG_TRY(exclusive_scan(ni, ntiles, istart, partial, true, s));
    hipLaunchKernelGGL(k_degrid_items, dim3(blocks_for(ntiles)), dim3(256), 0, s, offs, ni, istart,
                       ntiles, items);
    MG_HIP(hipGetLastError());
    if (convert) {
        const int64_t ng = (int64_t)kSubKernels * S * S;
        hipLaunchKernelGGL((k_degrid_convert<double, float2>), dim3(blocks_for(ng)), dim3(256), 0, s,
                           (const double*)job->gcf, (float2*)gtab, ng);
        MG_HIP(hipGetLastError());
    }
    DegridArgs a;
    a.uv = job->uv;
    a.gtab = convert ? gtab : job->gcf;
    a.grid = job->grid;
    a.out = vals_out;
    a.rec = rec;
    a.items = items;
    a.nitems = istart + ntiles;
    a.N = N;
    a.S = S;
    a.T = T;
    a.ntx = ntx;
    a.planar = o.layout == MIGRID_LAYOUT_PLANAR ? 1 : 0;
    const int wpc = o.wg_per_cu;
    MG_TRY(ctx->prof.begin(s));
    int rc;
    if (!f32)  // double visibilities, double grid, double products
        rc = run_degrid<double2, C2<double>, double, double, double>(a, V, special, lds, wpc, s);
    else if (in64)  // double visibilities, FP32 tiles
        rc = out64 ? run_degrid<float2, C2<float>, float, double, double>(a, V, special, lds, wpc, s)
                   : run_degrid<float2, C2<float>, float, double, float>(a, V, special, lds, wpc, s);
    else  // float visibilities: sums in the grid's type
        rc = out64 ? run_degrid<float2, C2<double>, double, float, double>(a, V, special, lds, wpc, s)
                   : run_degrid<float2, C2<float>, float, float, float>(a, V, special, lds, wpc, s);
    MG_TRY(rc);
    MG_TRY(ctx->prof.end(s));
    return MIGRID_OK;
}
