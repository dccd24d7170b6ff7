// Device-resident job tables written by kernels that take the entries by
// value: graph-safe (a captured launch owns its copy of the arguments, so a
// replay writes the same table) and without host staging buffers.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace migrid {

template <typename T>
struct TableChunk {
    static constexpr int kN = (int)(3072 / sizeof(T)) > 0 ? (int)(3072 / sizeof(T)) : 1;
    T v[kN];
};

template <typename T>
__global__ void k_put_table(T* __restrict__ dst, TableChunk<T> c, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = c.v[i];
}

// dst[0..n) = src[0..n) on stream s (one launch per TableChunk<T>::kN entries).
template <typename T>
int put_table(T* dst, const T* src, int n, hipStream_t s)
{
    constexpr int kN = TableChunk<T>::kN;
    for (int b = 0; b < n; b += kN) {
        TableChunk<T> c;
        const int m = n - b < kN ? n - b : kN;
        for (int i = 0; i < m; i++) c.v[i] = src[b + i];
        hipLaunchKernelGGL(k_put_table<T>, dim3(1), dim3(64), 0, s, dst + b, c, m);
    }
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

}  // namespace migrid
