// Device helpers shared by the gridding kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace migrid {

__device__ __forceinline__ double floor_t(double x) { return floor(x); }
__device__ __forceinline__ float floor_t(float x) { return floorf(x); }

// Native no-return hardware float atomics (global_atomic_add_f64 / _f32).
__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }

// Analytic GCF at tap (r, c) of sub-kernel (sx, sy):
// dx = (c - S/2 + 1) - sx/8, dy = (r - S/2 + 1) - sy/8, q = dx^2 + dy^2,
// g = exp(-q / (2 sigma^2)) * (cos(beta q) + i sin(beta q)), in FP64, the
// same expression the GCF table is filled with.
__device__ __forceinline__ void analytic_gcf(int c, int r, int S, int sx, int sy, double sigma,
                                             double beta, double& gr, double& gi)
{
    const double dx = (double)(c - S / 2 + 1) - sx / 8.0;
    const double dy = (double)(r - S / 2 + 1) - sy / 8.0;
    const double q = dx * dx + dy * dy;
    const double amp = exp(-q / (2.0 * sigma * sigma));
    double s, co;
    sincos(beta * q, &s, &co);
    gr = amp * co;
    gi = amp * s;
}

}  // namespace migrid
