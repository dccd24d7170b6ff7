// Host side of the C ABI: argument checks, device buffers, strategy dispatch.
//
// The drop-in entry points follow the reference host driver gridGPU
// (grid_gpu.cu:675-819): copy inputs to the device, zero the padded grid,
// run the strategy, copy the whole padded grid back.  Unlike the reference
// they return a status instead of printing and continuing
// (CUDA_CHECK_ERR, grid_gpu.cu:32-36), free everything they allocate
// (the reference leaks d_in, d_in_vals, d_gcf, :810-815), do not pin the
// caller's memory, and zero the full padded buffer for every type pair
// (the reference's scatter memset uses sizeof(CmplxType), :779).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "internal.hpp"

namespace migrid {

static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int fail_hip(hipError_t e, const char* what, const char* file, int line)
{
    const char* base = strrchr(file, '/');
    return fail(e == hipErrorOutOfMemory ? MIGRID_ENOMEM : MIGRID_EDEVICE,
                "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what,
                base ? base + 1 : file, line);
}

void clear_error() { g_err[0] = 0; }

bool default_planar(int strategy) { return strategy == MIGRID_WINDOW; }

int validate_shape(int64_t npts, int64_t N, int64_t S)
{
    if (npts < 0 || npts > INT32_MAX)
        return fail(MIGRID_EINVAL, "npts=%lld out of range [0, 2^31)", (long long)npts);
    if (N < 1 || N > 65536)
        return fail(MIGRID_EINVAL, "img_dim=%lld out of range [1, 65536]", (long long)N);
    if (S < 2 || S > kMaxSupport || (S & 1))
        return fail(MIGRID_EINVAL, "gcf_dim=%lld must be even and in [2, %d]", (long long)S,
                    kMaxSupport);
    return MIGRID_OK;
}

int Scratch::ensure(size_t want)
{
    if (want <= bytes) return MIGRID_OK;
    if (p) {
        MG_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
    }
    size_t round = (want + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
    MG_HIP(hipMalloc(&p, round));
    bytes = round;
    return MIGRID_OK;
}

void Scratch::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int Profiler::begin(hipStream_t s)
{
    if (!on) return MIGRID_OK;
    if (used + 2 > ev.size()) {
        for (int i = 0; i < 2; i++) {
            hipEvent_t e;
            MG_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
    }
    MG_HIP(hipEventRecord(ev[used], s));
    return MIGRID_OK;
}

int Profiler::end(hipStream_t s)
{
    if (!on) return MIGRID_OK;
    MG_HIP(hipEventRecord(ev[used + 1], s));
    used += 2;
    return MIGRID_OK;
}

void Profiler::release()
{
    for (auto e : ev) (void)hipEventDestroy(e);
    ev.clear();
    used = 0;
}

int launch_tiled(migrid_ctx* ctx, const GridArgs& a, const migrid_opts& o, hipStream_t s);

static size_t elem_bytes(int t) { return t == MIGRID_F64 ? 8 : 4; }

int device_job_args(const migrid_job* job, const migrid_opts* opts_in, GridArgs& a,
                    migrid_opts& o)
{
    if (opts_in)
        o = *opts_in;
    else
        migrid_opts_default(&o);
    if (!job) return fail(MIGRID_EINVAL, "job is NULL");
    MG_TRY(validate_shape(job->npts, job->img_dim, job->gcf_dim));
    if (job->in_type != MIGRID_F64 && job->in_type != MIGRID_F32)
        return fail(MIGRID_EINVAL, "in_type=%d", job->in_type);
    if (job->out_type != MIGRID_F64 && job->out_type != MIGRID_F32)
        return fail(MIGRID_EINVAL, "out_type=%d", job->out_type);
    if (job->in_type == MIGRID_F64 && job->out_type == MIGRID_F32 &&
        !(o.strategy == MIGRID_TILED && o.arith != MIGRID_ARITH_NATIVE))
        return fail(MIGRID_ENOTSUP, "double inputs with a float grid: the reference has no such "
                                    "instantiation (grid_gpu.cu:820-825); the tiled strategy "
                                    "offers it with FP32 arithmetic");
    if (!job->grid) return fail(MIGRID_EINVAL, "grid is NULL");
    if (job->npts > 0 && (!job->uv || !job->vals))
        return fail(MIGRID_EINVAL, "uv/vals NULL with npts=%lld", (long long)job->npts);
    if (o.gcf_mode != MIGRID_GCF_ANALYTIC && !job->gcf)
        return fail(MIGRID_EINVAL, "gcf is NULL (needed unless gcf_mode is ANALYTIC)");
    if (o.strategy < MIGRID_SCATTER || o.strategy > MIGRID_TILED)
        return fail(MIGRID_EINVAL, "strategy=%d", o.strategy);
    if (o.gcf_mode < MIGRID_GCF_TABLE || o.gcf_mode > MIGRID_GCF_ANALYTIC)
        return fail(MIGRID_EINVAL, "gcf_mode=%d", o.gcf_mode);

    a.uv = job->uv;
    a.vals = job->vals;
    a.gcf = job->gcf;
    a.grid = job->grid;
    a.grid_lo = job->grid_lo;
    a.npts = job->npts;
    a.N = (int)job->img_dim;
    a.S = (int)job->gcf_dim;
    a.in_type = job->in_type;
    a.out_type = job->out_type;
    a.planar = o.layout == MIGRID_LAYOUT_PLANAR ||
               (o.layout == MIGRID_LAYOUT_DEFAULT && default_planar(o.strategy));
    a.gcf_mode = o.gcf_mode;
    a.debug = o.debug_coverage != 0;
    a.sigma = o.gcf_sigma > 0 ? o.gcf_sigma : a.S / 6.0;
    a.beta = o.gcf_beta;

    return MIGRID_OK;
}

// Grid one job whose buffers are already on the device.
static int grid_device(migrid_ctx* ctx, const migrid_job* job, const migrid_opts* opts_in,
                       hipStream_t s)
{
    GridArgs a;
    migrid_opts o;
    MG_TRY(device_job_args(job, opts_in, a, o));
    if (o.strategy == MIGRID_SCATTER) {
        if (!o.accumulate) {
            size_t gbytes = (size_t)a.N * a.N * 2 * elem_bytes(a.out_type);
            MG_TRY(launch_zero(a.grid, gbytes, s));
            if (a.grid_lo) MG_TRY(launch_zero(a.grid_lo, gbytes, s));
        }
        if (ctx) MG_TRY(ctx->prof.begin(s));
        MG_TRY(launch_scatter(a, s));
        if (ctx) MG_TRY(ctx->prof.end(s));
        return MIGRID_OK;
    }
    // the tiled strategies clear a fresh grid themselves (tiles nobody stores)
    return launch_tiled(ctx, a, o, s);
}

// Shared body of the host entry points.  cpu_form: gridCPU semantics
// (unpadded planar base, 2*N*N scalars) instead of gridGPU's padded buffer.
static int grid_host(void* out, int out_type, const void* in, const void* vals, const void* gcf,
                     int in_type, size_t npts, size_t N, size_t S, const migrid_opts* opts_in,
                     bool cpu_form, bool accumulate)
{
    clear_error();
    MG_TRY(validate_shape((int64_t)npts, (int64_t)N, (int64_t)S));
    migrid_opts o;
    if (opts_in)
        o = *opts_in;
    else
        migrid_opts_default(&o);
    if (cpu_form) {
        o.layout = MIGRID_LAYOUT_PLANAR;
        o.accumulate = accumulate ? 1 : 0;
    }
    if (!out) return fail(MIGRID_EINVAL, "out is NULL");
    if (npts && (!in || !vals)) return fail(MIGRID_EINVAL, "in/in_vals is NULL");
    if (!gcf && o.gcf_mode != MIGRID_GCF_ANALYTIC) return fail(MIGRID_EINVAL, "gcf is NULL");

    const size_t ce_out = 2 * elem_bytes(out_type);  // complex element bytes
    const size_t ce_in = 2 * elem_bytes(in_type);
    const size_t out_elems = cpu_form ? N * N : migrid_padded_elems(N, S);
    const size_t base_off = cpu_form ? 0 : N * S + S;  // grid_gpu.cu:732
    const size_t out_bytes = out_elems * ce_out;

    int rc = MIGRID_OK;
    void *d_out = nullptr, *d_uv = nullptr, *d_vals = nullptr, *d_gcf = nullptr;
    hipStream_t s = nullptr;
    migrid_ctx* ctx = nullptr;
    auto cleanup = [&]() {
        if (s) (void)hipStreamSynchronize(s);
        if (d_out) (void)hipFree(d_out);
        if (d_uv) (void)hipFree(d_uv);
        if (d_vals) (void)hipFree(d_vals);
        if (d_gcf) (void)hipFree(d_gcf);
        if (ctx) migrid_ctx_destroy(ctx);
        if (s) (void)hipStreamDestroy(s);
    };
#define MG_STEP(call)                                                          \
    do {                                                                       \
        rc = (call);                                                           \
        if (rc != MIGRID_OK) {                                                 \
            cleanup();                                                         \
            return rc;                                                         \
        }                                                                      \
    } while (0)
#define MG_HSTEP(call)                                                         \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            rc = fail_hip(e_, #call, __FILE__, __LINE__);                      \
            cleanup();                                                         \
            return rc;                                                         \
        }                                                                      \
    } while (0)

    MG_HSTEP(hipStreamCreate(&s));
    int dev = 0;
    MG_HSTEP(hipGetDevice(&dev));
    MG_STEP(migrid_ctx_create(&ctx, dev));
    MG_HSTEP(hipMalloc(&d_out, out_bytes));
    if (npts) {
        MG_HSTEP(hipMalloc(&d_uv, npts * ce_in));
        MG_HSTEP(hipMalloc(&d_vals, npts * ce_in));
        MG_HSTEP(hipMemcpyAsync(d_uv, in, npts * ce_in, hipMemcpyHostToDevice, s));
        MG_HSTEP(hipMemcpyAsync(d_vals, vals, npts * ce_in, hipMemcpyHostToDevice, s));
    }
    if (gcf) {
        const size_t gb = (size_t)kSubKernels * S * S * ce_in;
        MG_HSTEP(hipMalloc(&d_gcf, gb));
        MG_HSTEP(hipMemcpyAsync(d_gcf, gcf, gb, hipMemcpyHostToDevice, s));
    }
    if (o.accumulate)
        MG_HSTEP(hipMemcpyAsync(d_out, out, out_bytes, hipMemcpyHostToDevice, s));
    else
        MG_HSTEP(hipMemsetAsync(d_out, 0, out_bytes, s));  // whole padded buffer (:755,772)

    migrid_job job;
    job.in_type = in_type;
    job.out_type = out_type;
    job.uv = d_uv;
    job.vals = d_vals;
    job.gcf = d_gcf;
    job.grid = (char*)d_out + base_off * ce_out;
    job.grid_lo = nullptr;
    job.npts = (int64_t)npts;
    job.img_dim = (int64_t)N;
    job.gcf_dim = (int64_t)S;
    migrid_opts o2 = o;
    o2.accumulate = 1;  // the buffer is already zeroed or loaded
    MG_STEP(grid_device(ctx, &job, &o2, s));
    MG_HSTEP(hipMemcpyAsync(out, d_out, out_bytes, hipMemcpyDeviceToHost, s));
    MG_HSTEP(hipStreamSynchronize(s));
    cleanup();
    return MIGRID_OK;
#undef MG_STEP
#undef MG_HSTEP
}

}  // namespace migrid

using namespace migrid;

extern "C" {

void migrid_opts_default(migrid_opts* o)
{
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->strategy = MIGRID_SCATTER;
    o->layout = MIGRID_LAYOUT_DEFAULT;
    o->gcf_mode = MIGRID_GCF_TABLE;
    o->arith = MIGRID_ARITH_NATIVE;
    o->gcf_beta = 0.05;
}

const char* migrid_last_error(void) { return g_err; }
const char* migrid_version(void) { return "migrid 0.1 (gfx950)"; }

size_t migrid_padded_elems(size_t N, size_t S) { return N * N + 2 * N * S + 2 * S; }

int migrid_gridGPU_zz(migrid_c128* out, const migrid_c128* in, const migrid_c128* in_vals,
                      size_t npts, size_t img_dim, const migrid_c128* gcf, size_t gcf_dim,
                      const migrid_opts* opts)
{
    return grid_host(out, MIGRID_F64, in, in_vals, gcf, MIGRID_F64, npts, img_dim, gcf_dim, opts,
                     false, false);
}

int migrid_gridGPU_zc(migrid_c128* out, const migrid_c64* in, const migrid_c64* in_vals,
                      size_t npts, size_t img_dim, const migrid_c64* gcf, size_t gcf_dim,
                      const migrid_opts* opts)
{
    return grid_host(out, MIGRID_F64, in, in_vals, gcf, MIGRID_F32, npts, img_dim, gcf_dim, opts,
                     false, false);
}

int migrid_gridGPU_cc(migrid_c64* out, const migrid_c64* in, const migrid_c64* in_vals,
                      size_t npts, size_t img_dim, const migrid_c64* gcf, size_t gcf_dim,
                      const migrid_opts* opts)
{
    return grid_host(out, MIGRID_F32, in, in_vals, gcf, MIGRID_F32, npts, img_dim, gcf_dim, opts,
                     false, false);
}

int migrid_gridCPU_zz(migrid_c128* out, const migrid_c128* in, const migrid_c128* in_vals,
                      size_t npts, size_t img_dim, const migrid_c128* gcf, size_t gcf_dim)
{
    return grid_host(out, MIGRID_F64, in, in_vals, gcf, MIGRID_F64, npts, img_dim, gcf_dim,
                     nullptr, true, false);
}

int migrid_gridCPU_pz_zz(migrid_c128* out, const migrid_c128* in, const migrid_c128* in_vals,
                         size_t npts, size_t img_dim, const migrid_c128* gcf, size_t gcf_dim)
{
    return grid_host(out, MIGRID_F64, in, in_vals, gcf, MIGRID_F64, npts, img_dim, gcf_dim,
                     nullptr, true, true);
}

int migrid_gridCPU_pz_zc(migrid_c128* out, const migrid_c64* in, const migrid_c64* in_vals,
                         size_t npts, size_t img_dim, const migrid_c64* gcf, size_t gcf_dim)
{
    return grid_host(out, MIGRID_F64, in, in_vals, gcf, MIGRID_F32, npts, img_dim, gcf_dim,
                     nullptr, true, true);
}

int migrid_gridCPU_pz_cc(migrid_c64* out, const migrid_c64* in, const migrid_c64* in_vals,
                         size_t npts, size_t img_dim, const migrid_c64* gcf, size_t gcf_dim)
{
    return grid_host(out, MIGRID_F32, in, in_vals, gcf, MIGRID_F32, npts, img_dim, gcf_dim,
                     nullptr, true, true);
}

int migrid_ctx_create(migrid_ctx** out, int device)
{
    clear_error();
    if (!out) return fail(MIGRID_EINVAL, "ctx out-pointer is NULL");
    int n = 0;
    MG_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(MIGRID_EINVAL, "device %d not in [0, %d)", device, n);
    migrid_ctx* c = new (std::nothrow) migrid_ctx;
    if (!c) return fail(MIGRID_ENOMEM, "host allocation of the context failed");
    c->device = device;
    *out = c;
    return MIGRID_OK;
}

int migrid_ctx_destroy(migrid_ctx* ctx)
{
    if (!ctx) return MIGRID_OK;
    ctx->bins.release();
    ctx->records.release();
    ctx->queue.release();
    ctx->misc.release();
    ctx->prof.release();
    ctx->fft.release();
    delete ctx;
    return MIGRID_OK;
}

int migrid_ctx_grid(migrid_ctx* ctx, const migrid_job* job, const migrid_opts* opts, void* stream)
{
    clear_error();
    if (!ctx) return fail(MIGRID_EINVAL, "ctx is NULL");
    MG_HIP(hipSetDevice(ctx->device));
    return grid_device(ctx, job, opts, (hipStream_t)stream);
}

int migrid_ctx_set_profiling(migrid_ctx* ctx, int enable)
{
    clear_error();
    if (!ctx) return fail(MIGRID_EINVAL, "ctx is NULL");
    ctx->prof.on = enable != 0;
    return MIGRID_OK;
}

int migrid_ctx_profile_read(migrid_ctx* ctx, double* ms_out, int cap)
{
    clear_error();
    if (!ctx) return fail(MIGRID_EINVAL, "ctx is NULL");
    const int n = (int)(ctx->prof.used / 2);
    for (int i = 0; i < n; i++) {
        MG_HIP(hipEventSynchronize(ctx->prof.ev[2 * i + 1]));
        float ms = 0.f;
        MG_HIP(hipEventElapsedTime(&ms, ctx->prof.ev[2 * i], ctx->prof.ev[2 * i + 1]));
        if (ms_out && i < cap) ms_out[i] = ms;
    }
    return n;
}

int migrid_ctx_profile_reset(migrid_ctx* ctx)
{
    clear_error();
    if (!ctx) return fail(MIGRID_EINVAL, "ctx is NULL");
    ctx->prof.used = 0;
    return MIGRID_OK;
}

}  // extern "C"
