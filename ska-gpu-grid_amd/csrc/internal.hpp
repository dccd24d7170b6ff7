// Internal declarations shared by the host API and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "migrid.h"

namespace migrid {

constexpr int kOversample = 8;  // GCF_GRID (Defines.h:60): 8 x 8 = 64 sub-kernels
constexpr int kSubKernels = kOversample * kOversample;
constexpr int kMaxSupport = 128;

// Error reporting: every failing path records a message for
// migrid_last_error() and returns the negative status.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail_hip(hipError_t e, const char* what, const char* file, int line);
void clear_error();

#define MG_HIP(call)                                                           \
    do {                                                                       \
        hipError_t mg_e_ = (call);                                             \
        if (mg_e_ != hipSuccess)                                               \
            return ::migrid::fail_hip(mg_e_, #call, __FILE__, __LINE__);       \
    } while (0)

#define MG_TRY(call)                                                           \
    do {                                                                       \
        int mg_rc_ = (call);                                                   \
        if (mg_rc_ != MIGRID_OK) return mg_rc_;                                \
    } while (0)

// Kernel-facing description of one gridding job (all pointers on device).
struct GridArgs {
    const void* uv = nullptr;
    const void* vals = nullptr;
    const void* gcf = nullptr;
    void* grid = nullptr;     // unpadded base
    void* grid_lo = nullptr;  // two-sum compensation grid (FP32 flush) or null
    int64_t npts = 0;
    int N = 0;
    int S = 0;
    int in_type = MIGRID_F64;
    int out_type = MIGRID_F64;
    bool planar = false;
    int gcf_mode = MIGRID_GCF_TABLE;
    bool debug = false;
    double sigma = 0.0;
    double beta = 0.0;
};

// Layout a strategy writes when opts->layout is DEFAULT.
bool default_planar(int strategy);

// Checks shared by every entry point (shape, support, pointers).
int validate_shape(int64_t npts, int64_t N, int64_t S);

// ---- launchers (each returns MIGRID_OK or a recorded error) ------------
int launch_scatter(const GridArgs& a, hipStream_t s);
// The tiled strategies of one job on caller-provided device scratch of cap
// bytes (bin sort, records, work queue); fails with MIGRID_ENOTSUP, before
// any launch, when the job needs more.
int launch_tiled_scratch(const GridArgs& a, const migrid_opts& o, void* scratch, size_t cap,
                         hipStream_t s);
// Checks a device job and its options and fills the kernel-facing
// arguments (the checks of migrid_ctx_grid).
int device_job_args(const migrid_job* job, const migrid_opts* opts_in, GridArgs& a,
                    migrid_opts& o);
int launch_zero(void* p, size_t bytes, hipStream_t s);  // fill kernel, graph-safe
int fill_u32(void* p, uint32_t v, int64_t n, hipStream_t s);

// Device scratch owned by a context.
struct Scratch {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want);
    void release();
};

// Event pairs around the main kernel of each call (profiling only).
struct Profiler {
    bool on = false;
    std::vector<hipEvent_t> ev;  // 2 per launch
    size_t used = 0;             // events in use
    int begin(hipStream_t s);    // before the main kernel
    int end(hipStream_t s);      // after it
    void release();
};

// Cached 2-D FFT plan of a context (k_image.hip).
struct FftPlan {
    void* handle = nullptr;  // hipfftHandle
    int64_t n = 0;
    int type = 0;
    void release();
};

}  // namespace migrid

struct migrid_ctx {
    migrid::Profiler prof;
    int device = 0;
    migrid::Scratch bins;     // bin-sort keys, permutation, offsets
    migrid::Scratch records;  // packed per-visibility records
    migrid::Scratch queue;    // work queue of the tiled gridder
    migrid::Scratch misc;
    migrid::FftPlan fft;
};
