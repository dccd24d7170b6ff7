// Measured ceilings of the two resources the gridding kernels are bound by.
//
// The reference measures only its own kernels (the "Gflops" print,
// grid_gpu.cu:789-791).  These probes measure what the hardware gives:
//
//  * LDS reads.  Every tap the tiled owner kernel executes reads one GCF
//    element from LDS (8 bytes for FP32 pairs, 16 for FP64 pairs), so LDS
//    read bandwidth bounds it beside FMA issue.  The probe runs the owner
//    kernel's occupancy (1024-thread workgroups, two per CU) and reads with
//    either a linear pattern (lane i reads element i of a row: no bank
//    conflicts) or the owner kernel's own pattern (an 8 x 8 lane block
//    reading a 2 x 2 set of elements at stride 8 with a given row pitch,
//    shifted by a wave-uniform offset each step, as a visibility shifts it).
//    Loaded values are kept alive without VALU work, so the probe is bound
//    by LDS alone.
//  * Global float atomics.  The plain scatter kernel issues 2 S^2 hardware
//    atomic adds per visibility.  The probe issues exactly those atomics --
//    the same visibilities, lanes, taps and addresses -- adding constants,
//    with no table reads and no products, so its rate is the atomic ceiling
//    at that address distribution.
#include <hip/hip_runtime.h>

#include <vector>

#include "device_common.hpp"
#include "internal.hpp"

namespace migrid {
namespace {

constexpr int kProbeLds = 64 * 1024;  // bytes per workgroup: two workgroups per CU
constexpr int kProbeThreads = 1024;

// Keeps loaded values alive without spending VALU instructions on them.
// Called after all eight reads of two steps are issued, so each waits only
// for its own value and the reads are in flight together.
__device__ __forceinline__ void use(const uint2& v) { asm volatile("" ::"v"(v.x), "v"(v.y)); }
__device__ __forceinline__ void use(const uint4& v)
{
    asm volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w));
}
template <typename E>
__device__ __forceinline__ void use8(const E (&v)[8])
{
#pragma unroll
    for (int i = 0; i < 8; i++) use(v[i]);
}

// P = 0: linear rows (lane i reads element i of a 64-element row).  P > 0:
// the owner kernel's lane block with a row pitch of P elements.  Four
// element reads per lane per step at a wave-uniform offset; two steps'
// eight reads are issued before any is waited for.
template <typename E, int P>
__global__ __launch_bounds__(kProbeThreads) void k_lds_probe(int iters, uint32_t* __restrict__ out,
                                                             unsigned long long* __restrict__ cycles)
{
    extern __shared__ __align__(16) unsigned char smem[];
    E* s = (E*)smem;
    constexpr int nel = kProbeLds / sizeof(E);
    static_assert(P == 0 || 30 + 30 * P < nel, "the lane block's reads stay inside the buffer");
    for (int i = threadIdx.x; i < kProbeLds / 4; i += blockDim.x)
        ((uint32_t*)smem)[i] = (uint32_t)i * 2654435761u;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lx = lane & 7, ly = lane >> 3;
    const int lb = P == 0 ? lane : ly * P + lx;
    const unsigned long long t0 = clock64();
    for (int it = 0; it < iters; it += 2) {
        E v[8];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int st = it + h;
            if (P == 0) {
                // four consecutive rows at a uniform base
                const int b = (w * 256 + st * 256) & (nel - 256);
                const E* r = s + b + lb;
                v[4 * h] = r[0];
                v[4 * h + 1] = r[64];
                v[4 * h + 2] = r[128];
                v[4 * h + 3] = r[192];
            } else {
                // a visibility's offset within the zero border: 0..15 rows and columns
                const int off = ((st * 13 + w) & 15) + ((st * 5 + 3 * w) & 15) * P;
                const E* r = s + off + lb;
                v[4 * h] = r[0];
                v[4 * h + 1] = r[8];
                v[4 * h + 2] = r[8 * P];
                v[4 * h + 3] = r[8 * P + 8];
            }
        }
        use8(v);
    }
    const unsigned long long t1 = clock64();
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)t1;
    if (threadIdx.x == 0) cycles[blockIdx.x] = t1 - t0;
}

// The plain scatter kernel's atomics (k_scatter.hip): one wave per
// visibility, lanes over min(S, 64) columns and 64 / S rows per pass, two
// atomic adds (re, im) per tap on the grid, interleaved.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_atomic_probe(const TI* __restrict__ uv, int64_t npts,
                                                      int N, int S, TO* __restrict__ out)
{
    const int h = S / 2;
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t wpb = blockDim.x >> 6;
    const int64_t nwaves = (int64_t)gridDim.x * wpb;
    const int CW = S < 64 ? S : 64;
    const int RP = 64 / CW;
    const int c0 = lane % CW, r0 = lane / CW;
    if (r0 >= RP) return;
    for (int64_t n = (int64_t)blockIdx.x * wpb + wid; n < npts; n += nwaves) {
        const int mx = (int)floor_t(uv[2 * n]), my = (int)floor_t(uv[2 * n + 1]);
        for (int r = r0; r < S; r += RP) {
            const int y = my + r - h + 1;
            if (y < 0 || y >= N) continue;
            for (int c = c0; c < S; c += CW) {
                const int x = mx + c - h + 1;
                if (x < 0 || x >= N) continue;
                const int64_t p = (int64_t)y * N + x;
                atomic_add(out + 2 * p, (TO)1);
                atomic_add(out + 2 * p + 1, (TO)1);
            }
        }
    }
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace
}  // namespace migrid

using namespace migrid;

// Pitches the probe is built for: every residue mod 32 in steps of 4, and
// two odd ones.
#define MG_PITCHES(X) X(96) X(100) X(104) X(108) X(112) X(116) X(120) X(124) X(97) X(99)

template <typename E, int P>
static int launch_lds_probe(int nwg, int iters, uint32_t* out, unsigned long long* cyc)
{
    MG_HIP(hipFuncSetAttribute((const void*)k_lds_probe<E, P>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, kProbeLds));
    hipLaunchKernelGGL((k_lds_probe<E, P>), dim3(nwg), dim3(kProbeThreads), kProbeLds, nullptr,
                       iters, out, cyc);
    MG_HIP(hipGetLastError());
    return MIGRID_OK;
}

template <typename E>
static int launch_lds_pitch(int pitch, int nwg, int iters, uint32_t* out, unsigned long long* cyc)
{
    if (pitch == 0) return launch_lds_probe<E, 0>(nwg, iters, out, cyc);
#define MG_CASE(P) \
    if (pitch == P) return launch_lds_probe<E, P>(nwg, iters, out, cyc);
    MG_PITCHES(MG_CASE)
#undef MG_CASE
    return fail(MIGRID_EINVAL, "pitch %d: 0 (linear) or one of 96..124 step 4, 97, 99", pitch);
}

extern "C" int migrid_ceiling_lds(int32_t elem_bytes, int32_t pitch, int32_t iters,
                                  migrid_ceiling* res)
{
    clear_error();
    if (!res) return fail(MIGRID_EINVAL, "result is NULL");
    if ((elem_bytes != 8 && elem_bytes != 16) || iters < 1)
        return fail(MIGRID_EINVAL, "elem_bytes 8 or 16, iters >= 1");
    int dev = 0, ncu = 0;
    MG_HIP(hipGetDevice(&dev));
    MG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int per_cu = 2, nwg = ncu * per_cu;
    uint32_t* d_out = nullptr;
    unsigned long long* d_cyc = nullptr;
    MG_HIP(hipMalloc(&d_out, (size_t)nwg * kProbeThreads * 4));
    MG_HIP(hipMalloc(&d_cyc, (size_t)nwg * 8));
    Timer t;
    auto run = [&]() -> int {
        MG_HIP(hipEventCreate(&t.a));
        MG_HIP(hipEventCreate(&t.b));
        for (int rep = 0; rep < 2; rep++) {  // the first launch warms up
            MG_HIP(hipEventRecord(t.a, nullptr));
            if (elem_bytes == 8)
                MG_TRY(launch_lds_pitch<uint2>(pitch, nwg, iters, d_out, d_cyc));
            else
                MG_TRY(launch_lds_pitch<uint4>(pitch, nwg, iters, d_out, d_cyc));
            MG_HIP(hipEventRecord(t.b, nullptr));
            MG_HIP(hipEventSynchronize(t.b));
        }
        float ms = 0.f;
        MG_HIP(hipEventElapsedTime(&ms, t.a, t.b));
        std::vector<unsigned long long> cyc(nwg);
        MG_HIP(hipMemcpy(cyc.data(), d_cyc, (size_t)nwg * 8, hipMemcpyDeviceToHost));
        double mean = 0;
        for (auto c : cyc) mean += (double)c;
        mean /= nwg;
        const double bytes = (double)nwg * kProbeThreads * ((iters + 1) / 2 * 2) * 4.0 * elem_bytes;
        res->ms = ms;
        res->amount = bytes;
        res->per_s = bytes / (ms * 1e-3);
        res->per_clk_cu = bytes / nwg * per_cu / mean;  // per s_memtime tick
        res->clock_ghz = mean / (ms * 1e-3) / 1e9;
        return MIGRID_OK;
    };
    const int rc = run();
    (void)hipFree(d_out);
    (void)hipFree(d_cyc);
    return rc;
}

extern "C" int migrid_ceiling_atomics(const migrid_job* job, migrid_ceiling* res)
{
    clear_error();
    if (!res || !job) return fail(MIGRID_EINVAL, "job / result is NULL");
    MG_TRY(validate_shape(job->npts, job->img_dim, job->gcf_dim));
    if (!job->uv || !job->grid) return fail(MIGRID_EINVAL, "uv / grid is NULL");
    if (job->in_type != MIGRID_F64 && job->in_type != MIGRID_F32)
        return fail(MIGRID_EINVAL, "in_type=%d", job->in_type);
    if (job->out_type != MIGRID_F64 && job->out_type != MIGRID_F32)
        return fail(MIGRID_EINVAL, "out_type=%d", job->out_type);
    const int N = (int)job->img_dim, S = (int)job->gcf_dim;
    const int64_t npts = job->npts;
    const int64_t want = (npts + 3) / 4;
    const int blocks = (int)(want < 8192 ? (want < 1 ? 1 : want) : 8192);  // as the scatter launch
    Timer t;
    MG_HIP(hipEventCreate(&t.a));
    MG_HIP(hipEventCreate(&t.b));
    MG_HIP(hipEventRecord(t.a, nullptr));
#define MG_ATOM(TI, TO)                                                                        \
    hipLaunchKernelGGL((k_atomic_probe<TI, TO>), dim3(blocks), dim3(256), 0, nullptr,          \
                       (const TI*)job->uv, npts, N, S, (TO*)job->grid)
    if (job->in_type == MIGRID_F64 && job->out_type == MIGRID_F64) MG_ATOM(double, double);
    else if (job->in_type == MIGRID_F64) MG_ATOM(double, float);
    else if (job->out_type == MIGRID_F64) MG_ATOM(float, double);
    else MG_ATOM(float, float);
#undef MG_ATOM
    MG_HIP(hipGetLastError());
    MG_HIP(hipEventRecord(t.b, nullptr));
    MG_HIP(hipEventSynchronize(t.b));
    float ms = 0.f;
    MG_HIP(hipEventElapsedTime(&ms, t.a, t.b));
    res->ms = ms;
    res->amount = 2.0 * S * S * (double)npts;  // every tap on the grid (the caller's inputs)
    res->per_s = res->amount / (ms * 1e-3);
    res->per_clk_cu = 0;
    res->clock_ghz = 0;
    return MIGRID_OK;
}
