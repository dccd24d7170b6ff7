// Input side of the gridder: physical (u, v, w) to grid pixels, and the
// grouping of visibilities by a per-visibility GCF index.
//
// The reference keeps a per-visibility GCF index (vis.gcfinx, vis.h:17, set
// per baseline by its HDF5 reader, hdf52struct.cpp:24,116) that no kernel
// reads, and a device conversion convert() (grid_gpu.cu:45-51) that maps a
// normalised coordinate pin in [-1, 1) to (pin + 1) * asize and splits it
// into main pixel and sub-pixel -- the split the gridder's key computation
// does for non-negative coordinates.  Here the conversion starts from metres
// and channel frequencies, and the index selects one of several GCF tables
// (w-planes) at gridding time.
#include <hip/hip_runtime.h>

#include "bin.hpp"
#include "internal.hpp"

namespace migrid {
namespace {

constexpr double kLightSpeed = 299792458.0;  // m/s

// Visibility n = row * nchan + chan.  Written without contraction so the
// result is the same double arithmetic as the CPU restatement.
__global__ __launch_bounds__(256) void k_uvw_to_pixels(const double* __restrict__ uvw, int64_t nrow,
                                                       const double* __restrict__ freq, int nchan,
                                                       double N, double cell, int nplanes,
                                                       double w_max, double* __restrict__ uv,
                                                       int32_t* __restrict__ gidx)
{
#pragma clang fp contract(off)
    const int64_t total = nrow * nchan;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < total;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = n / nchan;
        const int ch = (int)(n - r * nchan);
        const double sc = freq[ch] / kLightSpeed;
        uv[2 * n] = 0.5 * N + uvw[3 * r] * sc * N * cell;
        uv[2 * n + 1] = 0.5 * N + uvw[3 * r + 1] * sc * N * cell;
        if (nplanes > 0 && gidx) {
            const double wl = fabs(uvw[3 * r + 2] * sc);
            const double g = floor(wl / w_max * (double)nplanes);
            gidx[n] = g < (double)(nplanes - 1) ? (int32_t)g : nplanes - 1;
        }
    }
}
// Sort keys of the grouping: the GCF index, or ngcf for an index out of range.
__global__ __launch_bounds__(256) void k_index_keys(const int32_t* __restrict__ idx, int64_t n,
                                                    int ngcf, uint32_t* __restrict__ keys)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t g = idx[i];
        keys[i] = (g >= 0 && g < ngcf) ? (uint32_t)g : (uint32_t)ngcf;
    }
}

// The visibilities in group order: out[p] = in[perm[p]] for uv and values.
template <typename TI>
__global__ __launch_bounds__(256) void k_gather_groups(const int32_t* __restrict__ perm, int64_t n,
                                                       const TI* __restrict__ uv,
                                                       const TI* __restrict__ vals,
                                                       TI* __restrict__ uv_out,
                                                       TI* __restrict__ vals_out)
{
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = perm[p];
        uv_out[2 * p] = uv[2 * i];
        uv_out[2 * p + 1] = uv[2 * i + 1];
        vals_out[2 * p] = vals[2 * i];
        vals_out[2 * p + 1] = vals[2 * i + 1];
    }
}

int blocks_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : b > 8192 ? 8192 : b);
}

}  // namespace

// Groups the visibilities of one job by GCF index: uv_out/vals_out receive
// them in group order (stable within a group) and offsets[0..ngcf] the group
// starts (offsets[ngcf] = the first visibility with an index out of range).
// scratch: bin_scratch_bytes(npts, ngcf + 1) bytes; perm: npts ints.
int group_by_index(const int32_t* idx, int64_t npts, int ngcf, int in_type, const void* uv,
                   const void* vals, void* uv_out, void* vals_out, void* scratch, int32_t* perm,
                   int32_t* offsets, hipStream_t s)
{
    if (npts > 0) {
        hipLaunchKernelGGL(k_index_keys, dim3(blocks_for(npts)), dim3(256), 0, s, idx, npts, ngcf,
                           sort_key_buffer(scratch));
        MG_HIP(hipGetLastError());
    }
    MG_TRY(sort_keys(npts, ngcf + 1, scratch, perm, offsets, s));
    if (npts > 0) {
        if (in_type == MIGRID_F64)
            hipLaunchKernelGGL(k_gather_groups<double>, dim3(blocks_for(npts)), dim3(256), 0, s,
                               perm, npts, (const double*)uv, (const double*)vals,
                               (double*)uv_out, (double*)vals_out);
        else
            hipLaunchKernelGGL(k_gather_groups<float>, dim3(blocks_for(npts)), dim3(256), 0, s,
                               perm, npts, (const float*)uv, (const float*)vals, (float*)uv_out,
                               (float*)vals_out);
        MG_HIP(hipGetLastError());
    }
    return MIGRID_OK;
}

}  // namespace migrid

using namespace migrid;

extern "C" int migrid_uvw_to_pixels(const double* uvw, int64_t nrow, const double* freq,
                                    int32_t nchan, int64_t img_dim, double cell_rad,
                                    int32_t nplanes, double w_max, double* uv_out,
                                    int32_t* gcf_index_out, void* stream)
{
    clear_error();
    if (nrow < 0 || nchan < 1 || img_dim < 1) return fail(MIGRID_EINVAL, "nrow/nchan/img_dim");
    if (!(cell_rad > 0.0)) return fail(MIGRID_EINVAL, "cell_rad=%g must be > 0", cell_rad);
    if (nplanes < 0 || (nplanes > 0 && (!(w_max > 0.0) || !gcf_index_out)))
        return fail(MIGRID_EINVAL, "nplanes=%d needs w_max > 0 and gcf_index_out", nplanes);
    if (nrow > 0 && (!uvw || !freq || !uv_out)) return fail(MIGRID_EINVAL, "NULL buffer");
    if (nrow == 0) return MIGRID_OK;
    const int64_t total = nrow * (int64_t)nchan;
    hipLaunchKernelGGL(k_uvw_to_pixels, dim3(blocks_for(total)), dim3(256), 0, (hipStream_t)stream,
                       uvw, nrow, freq, nchan, (double)img_dim, cell_rad, nplanes, w_max, uv_out,
                       nplanes > 0 ? gcf_index_out : nullptr);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}
