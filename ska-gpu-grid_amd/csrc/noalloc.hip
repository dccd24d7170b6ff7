// Gridding that allocates nothing: migrid_grid_noalloc.
//
// The same work as migrid_ctx_grid (the same checks, plan and kernels), but
// the tiled strategies' device scratch -- bin-sort keys, permutation and CSR
// offsets, packed records, per-tile counts, the work queue and, for gather
// and window, the partial tiles of the ordered reduction -- lives in the
// file-scope array `scratch` below instead of buffers a context allocates.
// The plain scatter strategy needs no scratch.  No hipMalloc, hipFree or
// other allocating HIP routine is called, so callers that cannot allocate
// device memory can grid.
//
// Every GPU the library runs on holds its own copy of the array until the
// process ends.  Every call on one GPU shares that copy: calls that overlap
// on one GPU (on different streams or host threads) race.  The passes clear
// what must start at zero (queue state, bin histograms) with fill kernels
// on each call, so the previous call's contents do no harm.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"

namespace migrid {
namespace {

constexpr size_t kScratchBytes = MIGRID_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// In host code this name is a stand-in for the device array, not its device
// address: host code uses the address looked up per device below.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int scratch_ptr(int dev, void** out)
{
    if (dev < 0 || dev >= kMaxDevices)
        return fail(MIGRID_ENOTSUP, "device %d: the allocation-free entry point serves devices "
                                    "0..%d", dev, kMaxDevices - 1);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        MG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(MIGRID_EDEVICE, "no device address for the scratch array");
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return MIGRID_OK;
}

}  // namespace
}  // namespace migrid

using namespace migrid;

extern "C" int migrid_grid_noalloc(const migrid_job* job, const migrid_opts* opts, void* stream)
{
    clear_error();
    GridArgs a;
    migrid_opts o;
    MG_TRY(device_job_args(job, opts, a, o));
    const hipStream_t s = (hipStream_t)stream;
    if (o.strategy == MIGRID_SCATTER) {  // one atomic per tap, no scratch
        if (!o.accumulate) {
            const size_t gbytes = (size_t)a.N * a.N * 2 * (a.out_type == MIGRID_F64 ? 8 : 4);
            MG_TRY(launch_zero(a.grid, gbytes, s));
            if (a.grid_lo) MG_TRY(launch_zero(a.grid_lo, gbytes, s));
        }
        return launch_scatter(a, s);
    }
    int dev = 0;
    MG_HIP(hipGetDevice(&dev));
    void* p = nullptr;
    MG_TRY(scratch_ptr(dev, &p));
    return launch_tiled_scratch(a, o, p, kScratchBytes, s);
}
