// Degridding: the adjoint of gridding (grid -> visibilities).  No counterpart
// in the reference (the step after its path, for prediction / major cycles).
//
//   vals[n] = sum over the in-grid taps of visibility n of
//             conj(G[k][r][c]) * grid[main_y - S/2 + 1 + r][main_x - S/2 + 1 + c]
//
// with the gridder's index math, so <grid(v), g> = <v, degrid(g)>.
//
// MI355X layout: a visibility belongs to the tile that holds its window's
// corner, q = floor(uv) + S/2 binned by the tile edge T (the gridder's q).
// One workgroup per non-empty tile stages the tile and its (S - 1)-pixel halo
// of the grid in LDS once (zeros off the grid), then each wave takes whole
// visibilities: its lanes walk the sub-kernel's columns, reading a GCF row
// (coalesced, cache-resident: the tile's visibilities are sorted by sub-kernel)
// and the matching staged grid row, and a 64-lane reduction gives the value.
// Every visibility is computed by exactly one wave, so no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bin.hpp"
#include "device_common.hpp"
#include "internal.hpp"

namespace migrid {
namespace {

struct DegridArgs {
    const void* uv;
    const void* gtab;  // 64 * S^2 complex in the product type
    const void* grid;
    void* out;
    const int32_t* perm;
    const int32_t* offs;  // CSR over (tile, k) keys
    int N, S, T, ntx;
    int planar;
};

template <typename T>
struct C2 {
    T re, im;
};

__device__ __forceinline__ double wave_sum(double x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
__device__ __forceinline__ float wave_sum(float x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
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

// G: coefficient pair type, GS: staged grid pair type, TA: accumulation.
template <typename G, typename GS, typename TA, typename TI, typename TO>
__global__ __launch_bounds__(1024) void k_degrid(DegridArgs a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    GS* tile = (GS*)smem_raw;
    const int t = blockIdx.x;
    const int32_t r0 = a.offs[(int64_t)t * kSubKernels];
    const int32_t r1 = a.offs[(int64_t)(t + 1) * kSubKernels];
    if (r0 >= r1) return;
    const int N = a.N, S = a.S, T = a.T;
    const int tx = t % a.ntx, ty = t / a.ntx;
    const int X0 = tx * T - (S - 1), Y0 = ty * T - (S - 1);
    const int W = T + S - 1;
    const TO* __restrict__ grid = (const TO*)a.grid;
    const int64_t NN = (int64_t)N * N;
    for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
        const int yy = e / W, xx = e - yy * W;
        const int y = Y0 + yy, x = X0 + xx;
        GS g;
        g.re = 0;
        g.im = 0;
        if (x >= 0 && y >= 0 && x < N && y < N) {
            const int64_t p = (int64_t)y * N + x;
            g.re = (decltype(g.re))(a.planar ? grid[p] : grid[2 * p]);
            g.im = (decltype(g.im))(a.planar ? grid[NN + p] : grid[2 * p + 1]);
        }
        tile[e] = g;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int R = 64 / S;          // rows a pass covers (S <= 64)
    const int c = lane % S, rr = lane / S;
    const bool active = rr < R;
    const TI* __restrict__ uv = (const TI*)a.uv;
    const G* __restrict__ gtab = (const G*)a.gtab;
    TI* __restrict__ out = (TI*)a.out;
    for (int j = r0 + wave; j < r1; j += nw) {
        const int n = __builtin_amdgcn_readfirstlane(a.perm[j]);
        const TI u = uv[2 * (int64_t)n], v = uv[2 * (int64_t)n + 1];
        const TI fu = floor_t(u), fv = floor_t(v);
        const int k = kOversample * (int)floor_t((TI)kOversample * (v - fv)) +
                      (int)floor_t((TI)kOversample * (u - fu));
        const int xr = (int)fu + S / 2 - tx * T;  // q - tile origin, in [0, T)
        const int yr = (int)fv + S / 2 - ty * T;
        const G* __restrict__ gk = gtab + (int64_t)k * S * S;
        TA ar = 0, ai = 0;
        if (active)
            for (int r = rr; r < S; r += R) {
                const G g = gk[r * S + c];
                const GS x = tile[(yr + r) * W + xr + c];
                const TA gr = (TA)g.x, gi = (TA)g.y, xr_ = (TA)x.re, xi = (TA)x.im;
                ar += gr * xr_ + gi * xi;
                ai += gr * xi - gi * xr_;
            }
        ar = wave_sum(ar);
        ai = wave_sum(ai);
        if (lane == 0) {
            out[2 * (int64_t)n] = (TI)ar;
            out[2 * (int64_t)n + 1] = (TI)ai;
        }
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

template <typename G, typename GS, typename TA, typename TI, typename TO>
int run_degrid(const DegridArgs& a, size_t lds, int64_t ntiles, hipStream_t s)
{
    auto kern = k_degrid<G, GS, TA, TI, TO>;
    MG_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)ntiles), dim3(1024), lds, s, a);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
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
    int T = 64;
    while (T > 16 && (size_t)(T + S - 1) * (T + S - 1) * gsz > 160 * 1024) T /= 2;
    const size_t lds = (size_t)(T + S - 1) * (T + S - 1) * gsz;
    if (lds > 160 * 1024) return fail(MIGRID_ENOTSUP, "degridding tile does not fit in LDS");
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
    // arena: permutation, CSR offsets, the GCF in the product type
    const bool convert = f32 && in64;
    const size_t gtab_bytes = convert ? (size_t)kSubKernels * S * S * 8 : 0;
    const size_t need = up256((size_t)npts * 4) + up256((size_t)K * 4) + gtab_bytes;
    MG_TRY(ctx->records.ensure(need));
    char* ar = (char*)ctx->records.p;
    int32_t* perm = (int32_t*)ar;
    int32_t* offs = (int32_t*)(ar + up256((size_t)npts * 4));
    void* gtab = ar + up256((size_t)npts * 4) + up256((size_t)K * 4);
    MG_TRY(ctx->bins.ensure(bin_scratch_bytes(npts, K)));
    if (in64)
        hipLaunchKernelGGL(k_degrid_keys<double>, dim3(blocks_for(npts)), dim3(256), 0, s,
                           (const double*)job->uv, npts, N, S, T, ntx, sort_key_buffer(ctx->bins.p));
    else
        hipLaunchKernelGGL(k_degrid_keys<float>, dim3(blocks_for(npts)), dim3(256), 0, s,
                           (const float*)job->uv, npts, N, S, T, ntx, sort_key_buffer(ctx->bins.p));
    MG_HIP(hipGetLastError());
    MG_TRY(sort_keys(npts, K, ctx->bins.p, perm, offs, s));
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
    a.perm = perm;
    a.offs = offs;
    a.N = N;
    a.S = S;
    a.T = T;
    a.ntx = ntx;
    a.planar = o.layout == MIGRID_LAYOUT_PLANAR ? 1 : 0;
    MG_TRY(ctx->prof.begin(s));
    int rc;
    if (!f32)  // double visibilities, double grid, double products
        rc = run_degrid<double2, C2<double>, double, double, double>(a, lds, ntiles, s);
    else if (in64)  // double visibilities, FP32 tiles
        rc = out64 ? run_degrid<float2, C2<float>, float, double, double>(a, lds, ntiles, s)
                   : run_degrid<float2, C2<float>, float, double, float>(a, lds, ntiles, s);
    else  // float visibilities: sums in the grid's type
        rc = out64 ? run_degrid<float2, C2<double>, double, float, double>(a, lds, ntiles, s)
                   : run_degrid<float2, C2<float>, float, float, float>(a, lds, ntiles, s);
    MG_TRY(rc);
    MG_TRY(ctx->prof.end(s));
    return MIGRID_OK;
}
