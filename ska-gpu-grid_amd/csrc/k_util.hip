// Small fill/copy kernels.  Used instead of hipMemsetAsync/hipMemcpyAsync
// wherever a launch sequence may be captured into a hipGraph.
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace migrid {
namespace {

__global__ __launch_bounds__(256) void k_zero16(uint4* __restrict__ p, int64_t n)
{
    const uint4 z = make_uint4(0, 0, 0, 0);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z;
}

__global__ __launch_bounds__(256) void k_fill32(uint32_t* __restrict__ p, uint32_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int grid_blocks(int64_t n)
{
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return b < 1 ? 1 : (int)b;
}

}  // namespace

int fill_u32(void* p, uint32_t v, int64_t n, hipStream_t s)
{
    if (n <= 0) return MIGRID_OK;
    hipLaunchKernelGGL(k_fill32, dim3(grid_blocks(n)), dim3(256), 0, s, (uint32_t*)p, v, n);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

int launch_zero(void* p, size_t bytes, hipStream_t s)
{
    if (!bytes) return MIGRID_OK;
    if (bytes % 4) return fail(MIGRID_EINVAL, "zero fill of %zu bytes (not a multiple of 4)", bytes);
    const uintptr_t a = (uintptr_t)p;
    if (a % 16 == 0 && bytes % 16 == 0) {
        const int64_t n = (int64_t)(bytes / 16);
        hipLaunchKernelGGL(k_zero16, dim3(grid_blocks(n)), dim3(256), 0, s, (uint4*)p, n);
        MG_HIP(hipGetLastError());
        return MIGRID_OK;
    }
    return fill_u32(p, 0u, (int64_t)(bytes / 4), s);
}

}  // namespace migrid
