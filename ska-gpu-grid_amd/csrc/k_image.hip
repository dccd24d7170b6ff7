// Grid <-> image: the centred 2-D FFT that follows gridding (no counterpart
// in the reference, whose path ends at the grid).
//
//   to image:  img[y][x] = scale / (taper[x] taper[y]) *
//              sum_{v,u} grid[v][u] e^{+2 pi i ((u - N/2)(x - N/2) + (v - N/2)(y - N/2)) / N}
//   to grid:   the adjoint, grid = scale * sum_{y,x} img[y][x] / (taper[x] taper[y]) e^{-2 pi i ...}
//
// that is numpy's fftshift(ifft2(ifftshift(grid))) * N^2 (and
// fftshift(fft2(ifftshift(.)))), for even N.  The centring is a checkerboard
// sign: for even N, shifting by N/2 along an axis multiplies the transform by
// (-1)^k, and the two axes' constant factors (-1)^(N/2) cancel, so
//   centred(grid) = s . FFT2(s . grid),  s[y][x] = (-1)^(x + y).
// Since s . g transforms to the plain transform shifted by N/2 on both axes,
//   centred(grid)[y][x] = s[y][x] FFT2(grid)[(y + N/2) % N][(x + N/2) % N],
// and for an interleaved input with no taper to divide by beforehand the
// transform runs on the input directly and one pass after it swaps the
// quadrants in place with the sign, scale and taper.  Otherwise (planar
// input, or the to-grid taper) a pass before the transform applies the sign,
// the layout conversion and the taper, and the pass after it the sign and
// scale.  The transform itself is hipFFT (Z2Z / C2C, one cached plan per
// context).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include "internal.hpp"

namespace migrid {
namespace {

int fail_fft(hipfftResult r, const char* what)
{
    return fail(r == HIPFFT_ALLOC_FAILED ? MIGRID_ENOMEM : MIGRID_EDEVICE, "hipFFT error %d in %s",
                (int)r, what);
}

#define MG_FFT(call)                                                           \
    do {                                                                       \
        hipfftResult mg_r_ = (call);                                           \
        if (mg_r_ != HIPFFT_SUCCESS) return fail_fft(mg_r_, #call);            \
    } while (0)

// Before the transform: out[p] = s[p] * in[p] (/ taper in the to-grid
// direction), from the interleaved or planar input layout.  in may equal out
// (interleaved): each element is read before it is written by the same thread.
template <typename T>
__global__ __launch_bounds__(256) void k_fft_pre(const T* in, int planar, T* out, int64_t N,
                                                 const double* __restrict__ taper)
{
    const int64_t NN = N * N;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < NN;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = p / N, x = p - y * N;
        T re = planar ? in[p] : in[2 * p];
        T im = planar ? in[NN + p] : in[2 * p + 1];
        double f = ((x + y) & 1) ? -1.0 : 1.0;
        if (taper) f /= taper[x] * taper[y];
        re = (T)(f * (double)re);
        im = (T)(f * (double)im);
        out[2 * p] = re;
        out[2 * p + 1] = im;
    }
}

// After the transform, in place: x[p] *= s[p] * scale (/ taper in the
// to-image direction).
template <typename T>
__global__ __launch_bounds__(256) void k_fft_post(T* __restrict__ io, int64_t N, double scale,
                                                  const double* __restrict__ taper)
{
    const int64_t NN = N * N;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < NN;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = p / N, x = p - y * N;
        double f = ((x + y) & 1) ? -scale : scale;
        if (taper) f /= taper[x] * taper[y];
        io[2 * p] = (T)(f * (double)io[2 * p]);
        io[2 * p + 1] = (T)(f * (double)io[2 * p + 1]);
    }
}

// After a transform of the unsigned input, in place: the quadrant swap
// (y, x) <-> ((y + N/2) % N, (x + N/2) % N) with the sign (the same at both
// ends of a pair, N even), scale and taper (to-image direction).  One thread
// per pair: y < N/2.
template <typename T>
__global__ __launch_bounds__(256) void k_fft_post_swap(T* __restrict__ io, int64_t N, double scale,
                                                       const double* __restrict__ taper)
{
    const int64_t h = N / 2, half = h * N;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < half;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = p / N, x = p - y * N;
        const int64_t y2 = y + h, x2 = x < h ? x + h : x - h;
        const int64_t q = y2 * N + x2;
        const double f = ((x + y) & 1) ? -scale : scale;
        const double fp = taper ? f / (taper[x] * taper[y]) : f;
        const double fq = taper ? f / (taper[x2] * taper[y2]) : f;
        const T a_re = io[2 * p], a_im = io[2 * p + 1];
        const T b_re = io[2 * q], b_im = io[2 * q + 1];
        io[2 * p] = (T)(fp * (double)b_re);
        io[2 * p + 1] = (T)(fp * (double)b_im);
        io[2 * q] = (T)(fq * (double)a_re);
        io[2 * q + 1] = (T)(fq * (double)a_im);
    }
}

int blocks_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : b > 16384 ? 16384 : b);
}

int get_plan(migrid_ctx* ctx, int64_t N, int type, hipfftHandle* out)
{
    FftPlan& P = ctx->fft;
    if (!P.handle || P.n != N || P.type != type) {
        P.release();
        hipfftHandle h = nullptr;
        MG_FFT(hipfftCreate(&h));
        size_t work = 0;
        const hipfftResult r =
            hipfftMakePlan2d(h, (int)N, (int)N, type == MIGRID_F64 ? HIPFFT_Z2Z : HIPFFT_C2C, &work);
        if (r != HIPFFT_SUCCESS) {
            hipfftDestroy(h);
            return fail_fft(r, "hipfftMakePlan2d");
        }
        P.handle = (void*)h;
        P.n = N;
        P.type = type;
    }
    *out = (hipfftHandle)P.handle;
    return MIGRID_OK;
}

}  // namespace

void FftPlan::release()
{
    if (handle) hipfftDestroy((hipfftHandle)handle);
    handle = nullptr;
    n = 0;
    type = 0;
}

}  // namespace migrid

using namespace migrid;

extern "C" int migrid_ctx_fft(migrid_ctx* ctx, int32_t type, int32_t layout, const void* in,
                              void* out, int64_t img_dim, int32_t direction, double scale,
                              const double* taper, void* stream)
{
    clear_error();
    if (!ctx || !in || !out) return fail(MIGRID_EINVAL, "NULL argument");
    if (type != MIGRID_F64 && type != MIGRID_F32) return fail(MIGRID_EINVAL, "type %d", type);
    if (direction != MIGRID_FFT_TO_IMAGE && direction != MIGRID_FFT_TO_GRID)
        return fail(MIGRID_EINVAL, "direction %d", direction);
    if (layout != MIGRID_LAYOUT_DEFAULT && layout != MIGRID_LAYOUT_INTERLEAVED &&
        layout != MIGRID_LAYOUT_PLANAR)
        return fail(MIGRID_EINVAL, "layout %d", layout);
    const bool planar = layout == MIGRID_LAYOUT_PLANAR;
    if (img_dim < 2 || img_dim % 2 || img_dim > 65536)
        return fail(MIGRID_EINVAL, "img_dim=%lld: an even size in [2, 65536]", (long long)img_dim);
    if (planar && in == out)
        return fail(MIGRID_EINVAL, "a planar input cannot be transformed in place");
    const int64_t N = img_dim;
    MG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    hipfftHandle plan = nullptr;
    MG_TRY(get_plan(ctx, N, type, &plan));
    MG_FFT(hipfftSetStream(plan, s));
    const int nb = blocks_for(N * N);
    const double* pre_taper = direction == MIGRID_FFT_TO_GRID ? taper : nullptr;
    const double* post_taper = direction == MIGRID_FFT_TO_IMAGE ? taper : nullptr;
    const int dir = direction == MIGRID_FFT_TO_IMAGE ? HIPFFT_BACKWARD : HIPFFT_FORWARD;
    // the sign as a quadrant swap after the transform, unless a pass before
    // it is needed anyway
    const bool swap = !planar && !pre_taper;
    const int nbs = blocks_for(N * N / 2);
    MG_TRY(ctx->prof.begin(s));
    if (type == MIGRID_F64) {
        double* o = (double*)out;
        if (!swap) {
            hipLaunchKernelGGL(k_fft_pre<double>, dim3(nb), dim3(256), 0, s, (const double*)in,
                               planar ? 1 : 0, o, N, pre_taper);
            MG_HIP(hipGetLastError());
        }
        MG_FFT(hipfftExecZ2Z(plan, (hipfftDoubleComplex*)(swap ? in : out),
                             (hipfftDoubleComplex*)o, dir));
        if (swap)
            hipLaunchKernelGGL(k_fft_post_swap<double>, dim3(nbs), dim3(256), 0, s, o, N, scale,
                               post_taper);
        else
            hipLaunchKernelGGL(k_fft_post<double>, dim3(nb), dim3(256), 0, s, o, N, scale,
                               post_taper);
    } else {
        float* o = (float*)out;
        if (!swap) {
            hipLaunchKernelGGL(k_fft_pre<float>, dim3(nb), dim3(256), 0, s, (const float*)in,
                               planar ? 1 : 0, o, N, pre_taper);
            MG_HIP(hipGetLastError());
        }
        MG_FFT(hipfftExecC2C(plan, (hipfftComplex*)(swap ? in : out), (hipfftComplex*)o, dir));
        if (swap)
            hipLaunchKernelGGL(k_fft_post_swap<float>, dim3(nbs), dim3(256), 0, s, o, N, scale,
                               post_taper);
        else
            hipLaunchKernelGGL(k_fft_post<float>, dim3(nb), dim3(256), 0, s, o, N, scale,
                               post_taper);
    }
    MG_HIP(hipGetLastError());
    MG_TRY(ctx->prof.end(s));
    return MIGRID_OK;
}
