// Batched gridding: many jobs (channels / snapshots) of uneven sizes in one
// persistent owner-kernel launch over a shared device work queue, with an
// optional hipGraph capture of the whole launch sequence and an autotuner.
//
// The reference grids one job per gridGPU call (grid_gpu.cu:675-819); a batch
// plans all jobs into one device arena (records, bin-sort output, tables),
// then enqueues per-job grid clears, bin sorts, record packing and work-item
// generation, and finally a single k_owner launch whose workgroups take
// (job, tile, k-range) items from the queue until it is empty.  Every step is
// a kernel (no memset/memcpy nodes), so the sequence can be captured into a
// graph once and replayed for new data in the same buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "internal.hpp"
#include "tiled.hpp"

using namespace migrid;

struct migrid_batch {
    migrid_ctx* ctx = nullptr;
    migrid_opts opts;
    std::vector<GridArgs> args;
    TiledPlan plan;
    void* arena = nullptr;
    void* bins = nullptr;
    bool graph_ready = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace migrid {

// Grid arguments of one job (same checks as a single call).
int job_args(const migrid_job* j, const migrid_opts& o, GridArgs& a)
{
    MG_TRY(validate_shape(j->npts, j->img_dim, j->gcf_dim));
    if (!j->grid) return fail(MIGRID_EINVAL, "job grid is NULL");
    if (j->npts > 0 && (!j->uv || !j->vals)) return fail(MIGRID_EINVAL, "job uv/vals NULL");
    if (o.gcf_mode != MIGRID_GCF_ANALYTIC && !j->gcf) return fail(MIGRID_EINVAL, "job gcf NULL");
    if ((j->in_type != MIGRID_F64 && j->in_type != MIGRID_F32) ||
        (j->out_type != MIGRID_F64 && j->out_type != MIGRID_F32))
        return fail(MIGRID_EINVAL, "job types %d/%d", j->in_type, j->out_type);
    a = GridArgs();
    a.uv = j->uv;
    a.vals = j->vals;
    a.gcf = j->gcf;
    a.grid = j->grid;
    a.grid_lo = j->grid_lo;
    a.npts = j->npts;
    a.N = (int)j->img_dim;
    a.S = (int)j->gcf_dim;
    a.in_type = j->in_type;
    a.out_type = j->out_type;
    a.planar = o.layout == MIGRID_LAYOUT_PLANAR ||
               (o.layout == MIGRID_LAYOUT_DEFAULT && default_planar(o.strategy));
    a.gcf_mode = o.gcf_mode;
    a.debug = o.debug_coverage != 0;
    a.sigma = o.gcf_sigma > 0 ? o.gcf_sigma : a.S / 6.0;
    a.beta = o.gcf_beta;
    return MIGRID_OK;
}

// Clears (unless accumulating), prepares and runs one planned launch.
int run_plan(const TiledPlan& p, const std::vector<GridArgs>& args, const migrid_opts& o,
             char* arena, void* bins, Profiler* prof, hipStream_t s, int n_sorted = -1,
             const void* const* kept = nullptr)
{
    // fresh grids: the plan's own passes clear what the owner kernel does not
    // store, except grids shared by several jobs, cleared here once each
    if (p.shared && !o.accumulate) {
        std::vector<void*> done;
        for (const GridArgs& a : args) {
            const size_t gb = (size_t)a.N * a.N * 2 * (a.out_type == MIGRID_F64 ? 8 : 4);
            for (void* g : {a.grid, a.grid_lo}) {
                if (!g || std::find(done.begin(), done.end(), g) != done.end()) continue;
                MG_TRY(launch_zero(g, gb, s));
                done.push_back(g);
            }
        }
    }
    MG_TRY(tiled_prepare(p, arena, bins, s, n_sorted, kept));
    if (prof) MG_TRY(prof->begin(s));
    MG_TRY(tiled_execute(p, arena, s));
    if (prof) MG_TRY(prof->end(s));
    return MIGRID_OK;
}

}  // namespace migrid

extern "C" {

int migrid_batch_create(migrid_ctx* ctx, const migrid_job* jobs, int njobs,
                        const migrid_opts* opts_in, migrid_batch** out)
{
    clear_error();
    if (!ctx || !out || (!jobs && njobs > 0)) return fail(MIGRID_EINVAL, "NULL argument");
    if (njobs < 1) return fail(MIGRID_EINVAL, "njobs=%d", njobs);
    migrid_opts o;
    if (opts_in)
        o = *opts_in;
    else
        migrid_opts_default(&o);
    o.strategy = o.strategy == MIGRID_SCATTER ? MIGRID_TILED : o.strategy;
    MG_HIP(hipSetDevice(ctx->device));
    migrid_batch* b = new (std::nothrow) migrid_batch;
    if (!b) return fail(MIGRID_ENOMEM, "host allocation of the batch failed");
    b->ctx = ctx;
    b->opts = o;
    b->args.resize(njobs);
    int rc = MIGRID_OK;
    for (int j = 0; j < njobs && rc == MIGRID_OK; j++) rc = job_args(&jobs[j], o, b->args[j]);
    if (rc == MIGRID_OK) rc = tiled_plan(b->args.data(), njobs, o, b->plan);
    if (rc == MIGRID_OK) {
        hipError_t e = hipMalloc(&b->arena, b->plan.arena_bytes);
        if (e == hipSuccess) e = hipMalloc(&b->bins, std::max<size_t>(b->plan.bin_bytes, 256));
        if (e != hipSuccess) rc = fail_hip(e, "hipMalloc(batch arena)", __FILE__, __LINE__);
    }
    if (rc == MIGRID_OK) {
        // the job table is written once; replays reuse it
        std::vector<JobDesc> d;
        tiled_descs(b->plan, (char*)b->arena, d);
        rc = tiled_write_descs(b->plan, (char*)b->arena, d, nullptr);
        if (rc == MIGRID_OK) {
            hipError_t e = hipDeviceSynchronize();
            if (e != hipSuccess) rc = fail_hip(e, "job table", __FILE__, __LINE__);
        }
    }
    if (rc != MIGRID_OK) {
        migrid_batch_destroy(b);
        return rc;
    }
    *out = b;
    return MIGRID_OK;
}

int migrid_batch_destroy(migrid_batch* b)
{
    if (!b) return MIGRID_OK;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->arena) (void)hipFree(b->arena);
    if (b->bins) (void)hipFree(b->bins);
    delete b;
    return MIGRID_OK;
}

int migrid_batch_run(migrid_batch* b, void* stream)
{
    clear_error();
    if (!b) return fail(MIGRID_EINVAL, "batch is NULL");
    MG_HIP(hipSetDevice(b->ctx->device));
    return run_plan(b->plan, b->args, b->opts, (char*)b->arena, b->bins, &b->ctx->prof,
                    (hipStream_t)stream);
}

int migrid_batch_graph_launch(migrid_batch* b, void* stream)
{
    clear_error();
    if (!b) return fail(MIGRID_EINVAL, "batch is NULL");
    hipStream_t s = (hipStream_t)stream;
    MG_HIP(hipSetDevice(b->ctx->device));
    if (!b->graph_ready) {
        if (!s) return fail(MIGRID_EINVAL, "graph capture needs a non-null stream");
        MG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        int rc = run_plan(b->plan, b->args, b->opts, (char*)b->arena, b->bins, nullptr, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != MIGRID_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail_hip(e, "hipStreamEndCapture", __FILE__, __LINE__);
        b->graph = g;
        MG_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->graph_ready = true;
    }
    MG_TRY(b->ctx->prof.begin(s));
    MG_HIP(hipGraphLaunch(b->exec, s));
    MG_TRY(b->ctx->prof.end(s));
    return MIGRID_OK;
}

int migrid_batch_info(migrid_batch* b, int64_t* arena_bytes, int32_t* tile, int32_t* njobs)
{
    clear_error();
    if (!b) return fail(MIGRID_EINVAL, "batch is NULL");
    if (arena_bytes) *arena_bytes = (int64_t)(b->plan.arena_bytes + b->plan.bin_bytes);
    if (tile) *tile = b->plan.key.T;
    if (njobs) *njobs = (int32_t)b->args.size();
    return MIGRID_OK;
}

// ---- autotuner ------------------------------------------------------------------
// Times every (tile edge, work-item size) pair on the given jobs and keeps the
// fastest; results are cached per (S-set, N-set, arithmetic, npts decade) so a
// later call with the same shape class returns at once.
static std::mutex g_tune_mu;
static std::map<std::tuple<std::vector<int>, int, int, int, int>, std::pair<migrid_opts, double>>
    g_tune_cache;

int migrid_autotune(migrid_ctx* ctx, const migrid_job* jobs, int njobs, const migrid_opts* base,
                    migrid_opts* best, double* best_ms, void* stream)
{
    clear_error();
    if (!ctx || !jobs || njobs < 1 || !best) return fail(MIGRID_EINVAL, "NULL argument");
    migrid_opts o;
    if (base)
        o = *base;
    else
        migrid_opts_default(&o);
    o.strategy = MIGRID_TILED;
    std::vector<int> sizes;
    int64_t total = 0;
    for (int j = 0; j < njobs; j++) {
        sizes.push_back((int)jobs[j].gcf_dim * 100000 + (int)jobs[j].img_dim);
        total += jobs[j].npts;
    }
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    int decade = 0;
    for (int64_t t = total; t >= 10; t /= 10) decade++;
    const auto key = std::make_tuple(sizes, o.arith, jobs[0].in_type, jobs[0].out_type, decade);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            *best = it->second.first;
            if (best_ms) *best_ms = it->second.second;
            return MIGRID_OK;
        }
    }
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    MG_HIP(hipEventCreate(&e0));
    MG_HIP(hipEventCreate(&e1));
    // one timed run (after a warm-up) of the batch built with options c
    auto trial = [&](const migrid_opts& c, float* ms) -> int {
        migrid_batch* b = nullptr;
        int r = migrid_batch_create(ctx, jobs, njobs, &c, &b);
        if (r != MIGRID_OK) return r;
        r = migrid_batch_run(b, s);  // warm-up
        if (r == MIGRID_OK) {
            hipError_t e = hipEventRecord(e0, s);
            if (e == hipSuccess) r = migrid_batch_run(b, s);
            if (e == hipSuccess && r == MIGRID_OK) e = hipEventRecord(e1, s);
            if (e == hipSuccess && r == MIGRID_OK) e = hipEventSynchronize(e1);
            if (e == hipSuccess && r == MIGRID_OK) e = hipEventElapsedTime(ms, e0, e1);
            if (e != hipSuccess && r == MIGRID_OK) r = fail_hip(e, "autotune timing", __FILE__, __LINE__);
        }
        migrid_batch_destroy(b);
        return r;
    };
    const int tiles[] = {16, 32, 64};
    const int vpws[] = {64, 128, 256};  // visibilities per wavefront; an item holds 64 x that
    double bt = 1e30;
    migrid_opts bo = o;
    int rc = MIGRID_OK;
    for (int t : tiles) {
        for (int v : vpws) {
            migrid_opts c = o;
            c.tile = t;
            c.vis_per_wave = v * 64;
            float ms = 0.f;
            rc = trial(c, &ms);
            if (rc != MIGRID_OK) break;
            if (ms < bt) {
                bt = ms;
                bo = c;
            }
        }
        if (rc != MIGRID_OK) break;
    }
    // then the LDS split around the best shape: one or two workgroups sharing
    // a CU's LDS, and the staging threshold (a sub-kernel with fewer
    // visibilities for a tile is read from HBM rather than staged in LDS)
    if (rc == MIGRID_OK) {
        const migrid_opts shape = bo;
        for (int w : {1, 2}) {
            migrid_opts c = shape;
            c.wg_per_cu = w;
            float ms = 0.f;
            rc = trial(c, &ms);
            if (rc != MIGRID_OK) break;
            if (ms < bt) {
                bt = ms;
                bo = c;
            }
        }
        const migrid_opts split = bo;
        for (int m : {2, 16}) {
            if (rc != MIGRID_OK) break;
            migrid_opts c = split;
            c.stage_min = m;
            float ms = 0.f;
            rc = trial(c, &ms);
            if (rc != MIGRID_OK) break;
            if (ms < bt) {
                bt = ms;
                bo = c;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != MIGRID_OK) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = std::make_pair(bo, bt);
    }
    *best = bo;
    if (best_ms) *best_ms = bt;
    return MIGRID_OK;
}

}  // extern "C"

extern "C" int migrid_ctx_grid_indexed(migrid_ctx* ctx, const migrid_job* job, int32_t ngcf,
                                       const int32_t* gcf_index, const migrid_opts* opts_in,
                                       void* stream)
{
    clear_error();
    if (!ctx || !job) return fail(MIGRID_EINVAL, "NULL argument");
    migrid_opts o;
    if (opts_in)
        o = *opts_in;
    else
        migrid_opts_default(&o);
    if (o.strategy == MIGRID_SCATTER) o.strategy = MIGRID_TILED;
    if (o.gcf_mode != MIGRID_GCF_TABLE)
        return fail(MIGRID_ENOTSUP, "indexed gridding reads GCF tables (gcf_mode TABLE)");
    if (ngcf < 1 || ngcf >= (1 << 24)) return fail(MIGRID_EINVAL, "ngcf=%d", ngcf);
    if (job->npts > 0 && !gcf_index) return fail(MIGRID_EINVAL, "gcf_index is NULL");
    GridArgs whole;
    MG_TRY(job_args(job, o, whole));
    MG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = job->npts;
    const size_t ce = job->in_type == MIGRID_F64 ? 16 : 8;  // complex element
    const size_t gcf_elems = (size_t)kSubKernels * job->gcf_dim * job->gcf_dim;

    // group the visibilities by index (device), then read the group sizes
    const size_t vbytes = (size_t)(n > 0 ? n : 1) * ce;
    const size_t misc = 2 * ((vbytes + 255) & ~(size_t)255) + (((size_t)(n > 0 ? n : 1) * 4 + 255) & ~(size_t)255) +
                        ((size_t)(ngcf + 1) * 4 + 255);
    MG_TRY(ctx->misc.ensure(misc));
    char* m = (char*)ctx->misc.p;
    void* guv = m;
    void* gvals = m + ((vbytes + 255) & ~(size_t)255);
    int32_t* perm = (int32_t*)(m + 2 * ((vbytes + 255) & ~(size_t)255));
    int32_t* goffs = (int32_t*)((char*)perm + (((size_t)(n > 0 ? n : 1) * 4 + 255) & ~(size_t)255));
    MG_TRY(ctx->bins.ensure(bin_scratch_bytes(n, ngcf + 1)));
    MG_TRY(group_by_index(gcf_index, n, ngcf, job->in_type, job->uv, job->vals, guv, gvals,
                          ctx->bins.p, perm, goffs, s));
    std::vector<int32_t> h(ngcf + 1);
    MG_HIP(hipMemcpyAsync(h.data(), goffs, (ngcf + 1) * 4, hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
    if (h[ngcf] != n)
        return fail(MIGRID_EINVAL, "%lld visibilities have a GCF index outside [0, %d)",
                    (long long)(n - h[ngcf]), ngcf);

    // one job per non-empty group, all into the caller's grid
    std::vector<GridArgs> args;
    for (int g = 0; g < ngcf; g++) {
        const int64_t lo = h[g], hi = g + 1 < ngcf ? h[g + 1] : h[ngcf];
        if (hi <= lo) continue;
        GridArgs a = whole;
        a.uv = (const char*)guv + (size_t)lo * ce;
        a.vals = (const char*)gvals + (size_t)lo * ce;
        a.gcf = (const char*)job->gcf + (size_t)g * gcf_elems * ce;
        a.npts = hi - lo;
        args.push_back(a);
    }
    if (args.empty()) {  // no visibilities: a fresh grid is only cleared
        if (o.accumulate) return MIGRID_OK;
        const size_t gb = (size_t)whole.N * whole.N * 2 * (whole.out_type == MIGRID_F64 ? 8 : 4);
        MG_TRY(launch_zero(whole.grid, gb, s));
        if (whole.grid_lo) MG_TRY(launch_zero(whole.grid_lo, gb, s));
        return MIGRID_OK;
    }
    TiledPlan p;
    MG_TRY(tiled_plan(args.data(), (int)args.size(), o, p));
    MG_TRY(ctx->bins.ensure(p.bin_bytes));
    MG_TRY(ctx->records.ensure(p.arena_bytes));
    char* arena = (char*)ctx->records.p;
    std::vector<JobDesc> d;
    tiled_descs(p, arena, d);
    MG_TRY(tiled_write_descs(p, arena, d, s));
    return run_plan(p, args, o, arena, ctx->bins.p, &ctx->prof, s);
}
