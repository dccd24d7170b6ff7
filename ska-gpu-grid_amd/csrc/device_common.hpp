// Device helpers shared by the gridding kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

namespace migrid {

__device__ __forceinline__ double floor_t(double x) { return floor(x); }
__device__ __forceinline__ float floor_t(float x) { return floorf(x); }

// Bin key of a visibility: (k, q_y / H, q_x / H) over q = floor(coordinate)
// + S/2 and k = 8 sy + sx, sx = floor(8 (u - floor(u))); nkeys (the job's
// sentinel) when q leaves [0, N + S - 2].  q is formed in the input type
// before narrowing, so a coordinate far outside the grid cannot wrap into it.
template <typename TI>
__device__ __forceinline__ uint32_t bin_key(TI u, TI v, int N, int S, int H, int nb,
                                            uint32_t nkeys)
{
    const TI fu = floor_t(u), fv = floor_t(v);
    const int sx = (int)floor_t((TI)kOversample * (u - fu));
    const int sy = (int)floor_t((TI)kOversample * (v - fv));
    const TI qxf = fu + (TI)(S / 2), qyf = fv + (TI)(S / 2);
    const int qmax = N + S - 2;
    if (qxf >= 0 && qyf >= 0 && qxf <= (TI)qmax && qyf <= (TI)qmax) {
        const int qx = (int)qxf, qy = (int)qyf;
        return ((uint32_t)(kOversample * sy + sx) * nb + qy / H) * nb + qx / H;
    }
    return nkeys;
}

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
