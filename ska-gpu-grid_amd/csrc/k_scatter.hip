// Plain scatter gridder: one hardware atomic per tap.
//
// Counterpart of the reference's default strategy, grid_kernel
// (grid_gpu.cu:95-173) with atomicAddWrap (:53-90).  Re-designed for wave64:
// one wavefront owns one visibility at a time; its 64 lanes cover the
// sub-kernel columns (c) of min(S,64) taps and 64/S rows per pass, so the GCF
// row read is one coalesced load and each tap issues a native
// global_atomic_add_f64 / _f32 (no CAS loop; the grid lives in coarse-grained
// hipMalloc memory).  The visibility's uv and value are wave-uniform and come
// through scalar loads.
//
// Tap semantics (grid.cu:115-132, grid_gpu.cu:124-163): for the visibility at
// (u, v), main = floor, sub = floor(8*frac), k = 8*sub_y + sub_x, tap (r, c) in
// [0,S)^2 lands on pixel (main_x + c - S/2 + 1, main_y + r - S/2 + 1) with
// coefficient gcf[(k*S + r)*S + c]; taps off the grid are dropped.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "internal.hpp"

namespace migrid {
namespace {

template <typename TI, typename TO, int SC, bool PLANAR, int GM, bool DBG>
__global__ __launch_bounds__(256) void k_scatter(const TI* __restrict__ uv,
                                                 const TI* __restrict__ vals,
                                                 const TI* __restrict__ gcf,
                                                 TO* __restrict__ out, int64_t npts, int N,
                                                 int S_rt, double sigma, double beta)
{
    const int S = SC ? SC : S_rt;
    const int h = S / 2;
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t wpb = blockDim.x >> 6;
    const int64_t nwaves = (int64_t)gridDim.x * wpb;
    const int CW = S < 64 ? S : 64;  // columns per pass
    const int RP = 64 / CW;          // rows per pass
    const int c0 = lane % CW;
    const int r0 = lane / CW;
    const bool lane_on = r0 < RP;
    const int64_t NN = (int64_t)N * N;

    // COMPUTE_GCF constants (grid_gpu.cu:103-108): taken from the element
    // the reference's pre-offset table pointer addresses, S(S-1)/2 - 1.
    float p1 = 0.f, p2 = 0.f;
    if (GM == MIGRID_GCF_COMPAT) {
        const int64_t e = (int64_t)S * (S - 1) / 2 - 1;
        double T = (double)gcf[2 * e];
        double w = (double)gcf[2 * e + 1];
        p1 = (float)(2 * 3.1415926 * w);
        p2 = (float)(p1 * T);
    }

    for (int64_t n = (int64_t)blockIdx.x * wpb + wid; n < npts; n += nwaves) {
        const TI u = uv[2 * n], v = uv[2 * n + 1];
        const TI vr = vals[2 * n], vi = vals[2 * n + 1];
        const int mx = (int)floor_t(u), my = (int)floor_t(v);
        const int sx = (int)floor_t((TI)kOversample * (u - floor_t(u)));
        const int sy = (int)floor_t((TI)kOversample * (v - floor_t(v)));
        const int k = kOversample * sy + sx;
        const bool interior = mx - h + 1 >= 0 && mx + h < N && my - h + 1 >= 0 && my + h < N;

        TI cr = 0, ci = 0;  // constant tap value of the COMPAT form
        if (GM == MIGRID_GCF_COMPAT) {
            float xs = (float)(mx - u + sx * 1.0 / 8.0);
            float ys = (float)(my - v + sy * 1.0 / 8.0);
            xs *= xs;
            ys *= ys;
            float ph = p1 - p2 * sqrtf(xs + ys);
            float s_, c_;
            sincosf(ph, &s_, &c_);
            cr = (TI)s_;  // r2 = sin, i2 = cos (grid_gpu.cu:146-147)
            ci = (TI)c_;
        }
        if (!lane_on) continue;
        const TI* g = gcf + 2 * (int64_t)k * S * S;
        for (int r = r0; r < S; r += RP) {
            const int y = my + r - h + 1;
            if (!interior && (y < 0 || y >= N)) continue;
            for (int c = c0; c < S; c += CW) {
                const int x = mx + c - h + 1;
                if (!interior && (x < 0 || x >= N)) continue;
                TI gr, gi;
                if (GM == MIGRID_GCF_TABLE) {
                    const int64_t t = (int64_t)r * S + c;
                    gr = g[2 * t];
                    gi = g[2 * t + 1];
                } else if (GM == MIGRID_GCF_COMPAT) {
                    gr = cr;
                    gi = ci;
                } else {
                    double gr_, gi_;
                    analytic_gcf(c, r, S, sx, sy, sigma, beta, gr_, gi_);
                    gr = (TI)gr_;
                    gi = (TI)gi_;
                }
                TO pr, pi;
                if (DBG) {
                    pr = (TO)1;
                    pi = (TO)n;
                } else {
                    pr = (TO)(vr * gr - vi * gi);
                    pi = (TO)(vr * gi + gr * vi);
                }
                const int64_t p = (int64_t)y * N + x;
                if (PLANAR) {
                    atomic_add(out + p, pr);
                    atomic_add(out + NN + p, pi);
                } else {
                    atomic_add(out + 2 * p, pr);
                    atomic_add(out + 2 * p + 1, pi);
                }
            }
        }
    }
}

template <typename TI, typename TO, int SC, bool PLANAR, int GM, bool DBG>
void launch_t(const GridArgs& a, hipStream_t s)
{
    const int threads = 256;
    int64_t want = (a.npts + 3) / 4;
    int blocks = (int)(want < 8192 ? (want < 1 ? 1 : want) : 8192);
    hipLaunchKernelGGL((k_scatter<TI, TO, SC, PLANAR, GM, DBG>), dim3(blocks), dim3(threads), 0, s,
                       (const TI*)a.uv, (const TI*)a.vals, (const TI*)a.gcf, (TO*)a.grid, a.npts,
                       a.N, a.S, a.sigma, a.beta);
}

template <typename TI, typename TO, int SC, bool PLANAR, int GM>
void pick_dbg(const GridArgs& a, hipStream_t s)
{
    if (a.debug)
        launch_t<TI, TO, SC, PLANAR, GM, true>(a, s);
    else
        launch_t<TI, TO, SC, PLANAR, GM, false>(a, s);
}

template <typename TI, typename TO, int SC, bool PLANAR>
void pick_gm(const GridArgs& a, hipStream_t s)
{
    switch (a.gcf_mode) {
    case MIGRID_GCF_COMPAT: pick_dbg<TI, TO, SC, PLANAR, MIGRID_GCF_COMPAT>(a, s); break;
    case MIGRID_GCF_ANALYTIC: pick_dbg<TI, TO, SC, PLANAR, MIGRID_GCF_ANALYTIC>(a, s); break;
    default: pick_dbg<TI, TO, SC, PLANAR, MIGRID_GCF_TABLE>(a, s); break;
    }
}

template <typename TI, typename TO, int SC>
void pick_layout(const GridArgs& a, hipStream_t s)
{
    if (a.planar)
        pick_gm<TI, TO, SC, true>(a, s);
    else
        pick_gm<TI, TO, SC, false>(a, s);
}

template <typename TI, typename TO>
void pick_s(const GridArgs& a, hipStream_t s)
{
    switch (a.S) {
    case 64: pick_layout<TI, TO, 64>(a, s); break;
    case 32: pick_layout<TI, TO, 32>(a, s); break;
    case 16: pick_layout<TI, TO, 16>(a, s); break;
    case 8: pick_layout<TI, TO, 8>(a, s); break;
    default: pick_layout<TI, TO, 0>(a, s); break;
    }
}

}  // namespace

int launch_scatter(const GridArgs& a, hipStream_t s)
{
    if (a.npts == 0) return MIGRID_OK;
    if (a.in_type == MIGRID_F64 && a.out_type == MIGRID_F64)
        pick_s<double, double>(a, s);
    else if (a.in_type == MIGRID_F32 && a.out_type == MIGRID_F64)
        pick_s<float, double>(a, s);
    else if (a.in_type == MIGRID_F32 && a.out_type == MIGRID_F32)
        pick_s<float, float>(a, s);
    else
        return fail(MIGRID_ENOTSUP, "scatter: unsupported type pair in=%d out=%d", a.in_type,
                    a.out_type);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

}  // namespace migrid
