// Serving loop: continuous batching of gridding requests on one GPU.
//
// The reference has no serving layer (one blocking gridGPU call per job,
// grid_gpu.cu:675-819).  Here a worker thread owns
//   * a copy stream: zeroing of new resident grids and uploads of requests
//     submitted from host memory (already copied into pinned staging);
//   * max_inflight batch slots, each with its own stream, arena and bin
//     scratch: a batch is the per-job bin sort, record packing and work
//     items, then ONE persistent owner-kernel launch for all its jobs.
// Requests are admitted in arrival order while their device footprint (a new
// job's resident grid, uploaded inputs, work arena estimate) fits the HBM
// budget; everything admitted (at most one request per job id, so a batch
// never writes one grid twice) forms the next batch as soon as a slot is
// free, and goes to the device at once on that slot's stream.  Batches of
// different slots run concurrently -- the next batch's sort and packing fill
// the tail of the running persistent kernel, and its kernel takes the CUs the
// other one leaves -- while a batch that writes a grid waits (stream event)
// for the previous batch that wrote it, so requests for one grid accumulate
// in submission order.  Grids stay resident per job id until dropped, and so
// does a job's bin sort when its request asks to keep it: requests that
// repeat its coordinates then skip the sort and only pack their values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <set>
#include <thread>
#include <vector>

#include "internal.hpp"
#include "tiled.hpp"

using namespace migrid;

namespace migrid {
int job_args(const migrid_job* j, const migrid_opts& o, GridArgs& a);
int run_plan(const TiledPlan& p, const std::vector<GridArgs>& args, const migrid_opts& o,
             char* arena, void* bins, Profiler* prof, hipStream_t s, int n_sorted,
             const void* const* kept);
}  // namespace migrid

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t ticket = 0;
    int64_t job_id = 0;
    int64_t N = 0, S = 0, npts = 0;
    const void* uv = nullptr;  // device pointers used by the batch
    const void* vals = nullptr;
    const void* gcf = nullptr;
    bool host = false;
    int32_t flags = 0;       // MIGRID_SUBMIT_*
    bool counted = false;    // admitted: counted in its job's inflight
    void* pinned = nullptr;  // host submissions: uv | vals | gcf
    void* d_in = nullptr;    // their device copy
    size_t in_bytes = 0;
    size_t work = 0;             // admission estimate of its arena share
    hipEvent_t ready = nullptr;  // uploads (and grid zeroing) done
    Clock::time_point t_submit;
};

struct Job {
    void* grid = nullptr;
    size_t bytes = 0;
    int64_t N = 0;
    int64_t inflight = 0;  // requests admitted, not completed
    int readers = 0;       // host copies of the grid in progress
    int last_slot = -1;    // slot of the last batch that wrote the grid
    uint64_t last_seq = 0; // and its sequence number
    void* sort = nullptr;  // kept bin sort (tiled_keep_sort layout)
    size_t sort_bytes = 0;
    int64_t sort_npts = -1, sort_S = 0;  // what the kept sort is of
};

struct Done {
    int status;
    double latency_ms;
};

// One batch slot: its own arena and bin scratch, reused batch after batch.
struct Slot {
    std::vector<Request> reqs;
    hipStream_t stream = nullptr;
    uint64_t seq = 0;  // sequence number of the batch last launched here
    hipEvent_t t0 = nullptr, t1 = nullptr;
    void* arena = nullptr;
    size_t arena_bytes = 0;
    void* bins = nullptr;
    size_t bins_bytes = 0;
};

}  // namespace

struct migrid_server {
    migrid_server_cfg cfg;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;       // wakes the worker
    std::condition_variable done_cv;  // wakes waiters
    bool stop = false;
    std::deque<Request> pending;    // submitted, not admitted
    std::vector<Request> admitted;  // admitted, not yet in a batch
    std::vector<Slot> slots;
    std::vector<int> busy;  // slots on the GPU
    std::vector<int> idle;
    std::map<int64_t, Job> jobs;
    std::map<int64_t, Done> done;
    int64_t released_below = 1;  // statuses of tickets below this were released
    int64_t next_ticket = 1;
    size_t hbm_used = 0;   // resident grids, kept sorts and uploaded inputs
    size_t work_used = 0;  // admission estimates of admitted and running work
    hipStream_t s_copy = nullptr;
    uint64_t batch_seq = 0;
    hipEvent_t t_ref = nullptr;  // the origin of the batches' device intervals
    std::vector<std::pair<double, double>> intervals;  // completed batches, ms from t_ref
    migrid_server_stats st;
    char last_err[256] = "";
};

namespace {

size_t grid_bytes(const migrid_server_cfg& c, int64_t N)
{
    return (size_t)N * N * 2 * (c.out_type == MIGRID_F64 ? 8 : 4);
}

size_t input_bytes(const migrid_server_cfg& c, int64_t npts, int64_t S)
{
    const size_t e = c.in_type == MIGRID_F64 ? 16 : 8;
    return (size_t)npts * e * 2 + (size_t)kSubKernels * S * S * e;
}

// Arena share of one request: records (<= 32 B), sort keys/permutation and
// offsets, work items, converted GCF table.  An upper estimate for admission.
size_t work_bytes(int64_t npts, int64_t N, int64_t S)
{
    const int64_t nb = (N + S - 2) / 8 + 1;  // bins of 8 (64-pixel tiles)
    const int64_t tiles = (N + 15) / 16;
    return (size_t)npts * 48 + (size_t)kSubKernels * nb * nb * 8 + (size_t)tiles * tiles * 64 +
           (size_t)kSubKernels * S * S * 16 + (1 << 20);
}

void release_inputs(migrid_server* s, Request& r)
{
    if (r.d_in) {  // its batch has completed
        (void)hipFreeAsync(r.d_in, s->s_copy);
        s->hbm_used -= r.in_bytes;
        r.d_in = nullptr;
    }
    if (r.pinned) {
        (void)hipHostFree(r.pinned);
        r.pinned = nullptr;
    }
    if (r.ready) {
        (void)hipEventDestroy(r.ready);
        r.ready = nullptr;
    }
}

void finish(migrid_server* s, Request& r, int status)
{
    const double ms =
        std::chrono::duration<double, std::milli>(Clock::now() - r.t_submit).count();
    s->done[r.ticket] = Done{status, ms};
    if (status == MIGRID_OK) {
        s->st.completed++;
        s->st.vis_completed += r.npts;
    } else {
        s->st.failed++;
    }
    release_inputs(s, r);
    if (r.counted) {
        s->work_used -= r.work;
        auto it = s->jobs.find(r.job_id);
        if (it != s->jobs.end()) it->second.inflight--;
        r.counted = false;
    }
}

// Completed statuses kept for poll/wait.  Past kDoneKeep entries, the oldest
// are released down to half of it -- only tickets below every pending one,
// which have all completed, so a released ticket is never one still running.
constexpr size_t kDoneKeep = (size_t)1 << 16;

void prune_done(migrid_server* s)
{
    if (s->done.size() <= kDoneKeep) return;
    int64_t min_pending = s->next_ticket;
    for (const Request& r : s->pending) min_pending = std::min(min_pending, r.ticket);
    for (const Request& r : s->admitted) min_pending = std::min(min_pending, r.ticket);
    for (int b : s->busy)
        for (const Request& r : s->slots[b].reqs) min_pending = std::min(min_pending, r.ticket);
    while (s->done.size() > kDoneKeep / 2) {
        auto it = s->done.begin();
        if (it->first >= min_pending) break;
        s->released_below = it->first + 1;
        s->done.erase(it);
    }
}

size_t arenas_bytes(const migrid_server* s)
{
    size_t b = 0;
    for (const Slot& sl : s->slots) b += sl.arena_bytes + sl.bins_bytes;
    return b;
}

// Admission of pending requests in arrival order (worker thread, mu held).
void admit(migrid_server* s)
{
    const bool quiet = s->busy.empty() && s->admitted.empty();
    while (!s->pending.empty()) {
        Request& r = s->pending.front();
        auto jt = s->jobs.find(r.job_id);
        const bool new_job = jt == s->jobs.end();
        if (!new_job && jt->second.N != r.N) {
            Request bad = r;
            s->pending.pop_front();
            snprintf(s->last_err, sizeof s->last_err, "job %lld holds a %lld^2 grid, request has %lld",
                     (long long)r.job_id, (long long)jt->second.N, (long long)r.N);
            finish(s, bad, MIGRID_EINVAL);
            continue;
        }
        r.work = work_bytes(r.npts, r.N, r.S);
        const size_t need = (new_job ? grid_bytes(s->cfg, r.N) : 0) +
                            (r.host ? input_bytes(s->cfg, r.npts, r.S) : 0) + r.work;
        if (need > (size_t)s->cfg.hbm_budget) {  // can never fit
            Request bad = r;
            s->pending.pop_front();
            snprintf(s->last_err, sizeof s->last_err, "request needs %zu B, budget %lld B", need,
                     (long long)s->cfg.hbm_budget);
            finish(s, bad, MIGRID_EBUSY);
            continue;
        }
        // memory held: resident grids, uploads, and the larger of the arenas
        // already allocated and the estimates of the work in flight
        const size_t held = s->hbm_used + std::max(arenas_bytes(s), s->work_used);
        if (held + need > (size_t)s->cfg.hbm_budget) break;  // wait for memory
        hipError_t e = hipSuccess;
        if (new_job) {
            Job j;
            j.N = r.N;
            j.bytes = grid_bytes(s->cfg, r.N);
            e = hipMalloc(&j.grid, j.bytes);
            if (e == hipSuccess && launch_zero(j.grid, j.bytes, s->s_copy) != MIGRID_OK)
                e = hipErrorLaunchFailure;
            if (e != hipSuccess) {
                if (j.grid) (void)hipFree(j.grid);
                (void)hipGetLastError();
                if (e == hipErrorOutOfMemory && !quiet) break;  // frees up as batches end
                Request bad = r;
                s->pending.pop_front();
                snprintf(s->last_err, sizeof s->last_err, "grid allocation: %s", hipGetErrorString(e));
                finish(s, bad, e == hipErrorOutOfMemory ? MIGRID_ENOMEM : MIGRID_EDEVICE);
                continue;
            }
            s->hbm_used += j.bytes;
            jt = s->jobs.emplace(r.job_id, j).first;
        }
        if (r.host) {
            const size_t b = input_bytes(s->cfg, r.npts, r.S);
            e = hipMallocAsync(&r.d_in, b, s->s_copy);
            if (e == hipSuccess) {
                s->hbm_used += b;
                r.in_bytes = b;
                e = hipMemcpyAsync(r.d_in, r.pinned, b, hipMemcpyHostToDevice, s->s_copy);
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                if (e == hipErrorOutOfMemory && !quiet) break;
                Request bad = r;
                s->pending.pop_front();
                snprintf(s->last_err, sizeof s->last_err, "upload: %s", hipGetErrorString(e));
                finish(s, bad, e == hipErrorOutOfMemory ? MIGRID_ENOMEM : MIGRID_EDEVICE);
                continue;
            }
            const size_t eb = s->cfg.in_type == MIGRID_F64 ? 16 : 8;
            r.uv = r.d_in;
            r.vals = (char*)r.d_in + (size_t)r.npts * eb;
            r.gcf = (char*)r.d_in + (size_t)r.npts * eb * 2;
        }
        e = hipEventCreateWithFlags(&r.ready, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(r.ready, s->s_copy);
        if (e != hipSuccess) {
            Request bad = r;
            s->pending.pop_front();
            snprintf(s->last_err, sizeof s->last_err, "event: %s", hipGetErrorString(e));
            finish(s, bad, MIGRID_EDEVICE);
            continue;
        }
        jt->second.inflight++;
        r.counted = true;
        s->work_used += r.work;
        s->admitted.push_back(r);
        s->pending.pop_front();
    }
}

int ensure(void*& p, size_t& have, size_t want)
{
    if (want <= have) return MIGRID_OK;
    if (p) MG_HIP(hipFree(p));
    p = nullptr;
    have = 0;
    MG_HIP(hipMalloc(&p, want));
    have = want;
    return MIGRID_OK;
}

// Plans and enqueues one batch from the admitted requests into slot `si`
// (mu held).  The requests taken are in `take` on return; on failure the
// caller completes them with the error.
int launch_batch(migrid_server* s, int si, std::vector<Request>& take)
{
    std::vector<Request> keep;
    std::set<int64_t> ids;
    for (Request& r : s->admitted) {
        if ((int)take.size() < s->cfg.max_batch && !ids.count(r.job_id) &&
            s->jobs[r.job_id].readers == 0) {
            ids.insert(r.job_id);
            take.push_back(r);
        } else {
            keep.push_back(r);
        }
    }
    s->admitted = std::move(keep);
    // a request that reuses its job's kept sort must match it (the job's
    // earlier requests are already launched); one that does not fails alone
    {
        std::vector<Request> ok;
        for (Request& r : take) {
            const Job& j = s->jobs[r.job_id];
            if ((r.flags & MIGRID_SUBMIT_SAME_UV) && !(j.sort_npts == r.npts && j.sort_S == r.S)) {
                snprintf(s->last_err, sizeof s->last_err,
                         "job %lld keeps no sort of %lld visibilities at S=%lld", (long long)r.job_id,
                         (long long)r.npts, (long long)r.S);
                finish(s, r, MIGRID_EINVAL);
                continue;
            }
            ok.push_back(r);
        }
        take = std::move(ok);
    }
    if (take.empty()) return MIGRID_OK;
    // the jobs sorted in this batch first, those taking a kept sort last
    std::stable_partition(take.begin(), take.end(),
                          [](const Request& r) { return !(r.flags & MIGRID_SUBMIT_SAME_UV); });
    int n_sorted = 0;
    while (n_sorted < (int)take.size() && !(take[n_sorted].flags & MIGRID_SUBMIT_SAME_UV))
        n_sorted++;
    migrid_opts o = s->cfg.opts;
    o.strategy = MIGRID_TILED;
    o.accumulate = 1;  // resident grids accumulate
    std::vector<GridArgs> args(take.size());
    for (size_t i = 0; i < take.size(); i++) {
        const Request& r = take[i];
        migrid_job j;
        memset(&j, 0, sizeof j);
        j.in_type = s->cfg.in_type;
        j.out_type = s->cfg.out_type;
        j.uv = r.uv;
        j.vals = r.vals;
        j.gcf = r.gcf;
        j.grid = s->jobs[r.job_id].grid;
        j.npts = r.npts;
        j.img_dim = r.N;
        j.gcf_dim = r.S;
        MG_TRY(job_args(&j, o, args[i]));
    }
    TiledPlan p;
    MG_TRY(tiled_plan(args.data(), (int)args.size(), o, p));
    Slot& sl = s->slots[si];
    // the slot is idle: nothing on the device uses its buffers
    MG_TRY(ensure(sl.arena, sl.arena_bytes, p.arena_bytes));
    MG_TRY(ensure(sl.bins, sl.bins_bytes, std::max<size_t>(p.bin_bytes, 256)));
    // room for the sorts this batch keeps (a buffer too small is replaced;
    // hipFree waits for the device, so no batch still reads the old one)
    for (size_t i = 0; i < take.size(); i++) {
        if (!(take[i].flags & MIGRID_SUBMIT_KEEP_SORT) || (take[i].flags & MIGRID_SUBMIT_SAME_UV))
            continue;
        Job& j = s->jobs[take[i].job_id];
        const size_t b = tiled_sort_bytes(p.jobs[i]);
        if (j.sort_bytes < b) {
            if (j.sort) {
                MG_HIP(hipFree(j.sort));
                s->hbm_used -= j.sort_bytes;
                j.sort = nullptr;
                j.sort_bytes = 0;
            }
            MG_HIP(hipMalloc(&j.sort, b));
            j.sort_bytes = b;
            s->hbm_used += b;
        }
    }
    // wait for the uploads, and for the previous batch that wrote each grid
    // (it also wrote any kept sort this batch reads)
    for (const Request& r : take) {
        MG_HIP(hipStreamWaitEvent(sl.stream, r.ready, 0));
        const Job& j = s->jobs[r.job_id];
        if (j.last_slot >= 0 && j.last_slot != si && s->slots[j.last_slot].seq == j.last_seq)
            MG_HIP(hipStreamWaitEvent(sl.stream, s->slots[j.last_slot].t1, 0));
    }
    std::vector<const void*> kept;
    for (size_t i = n_sorted; i < take.size(); i++) kept.push_back(s->jobs[take[i].job_id].sort);
    MG_HIP(hipEventRecord(sl.t0, sl.stream));
    std::vector<JobDesc> d;
    tiled_descs(p, (char*)sl.arena, d);
    MG_TRY(tiled_write_descs(p, (char*)sl.arena, d, sl.stream));
    MG_TRY(run_plan(p, args, o, (char*)sl.arena, sl.bins, nullptr, sl.stream, n_sorted,
                    kept.data()));
    for (size_t i = 0; i < take.size(); i++) {
        if (!(take[i].flags & MIGRID_SUBMIT_KEEP_SORT) || (take[i].flags & MIGRID_SUBMIT_SAME_UV))
            continue;
        Job& j = s->jobs[take[i].job_id];
        MG_TRY(tiled_keep_sort(p, (const char*)sl.arena, (int)i, j.sort, sl.stream));
        j.sort_npts = take[i].npts;
        j.sort_S = take[i].S;
    }
    MG_HIP(hipEventRecord(sl.t1, sl.stream));
    sl.seq = ++s->batch_seq;
    for (const Request& r : take) {
        Job& j = s->jobs[r.job_id];
        j.last_slot = si;
        j.last_seq = sl.seq;
    }
    s->st.sorts_reused += (int64_t)take.size() - n_sorted;
    sl.reqs = std::move(take);
    take.clear();
    s->st.batches++;
    return MIGRID_OK;
}

// Completes every finished batch (mu held).  Returns true when at least one
// finished.
bool reap(migrid_server* s)
{
    bool any = false;
    for (size_t q = 0; q < s->busy.size();) {
        const int si = s->busy[q];
        Slot& sl = s->slots[si];
        const hipError_t e = hipEventQuery(sl.t1);
        if (e == hipErrorNotReady) {
            q++;
            continue;
        }
        int status = MIGRID_OK;
        if (e != hipSuccess) {
            status = MIGRID_EDEVICE;
            snprintf(s->last_err, sizeof s->last_err, "batch failed: %s", hipGetErrorString(e));
        } else {
            float ms = 0.f, a = 0.f, b = 0.f;
            if (hipEventElapsedTime(&ms, sl.t0, sl.t1) == hipSuccess) s->st.gpu_ms += ms;
            if (hipEventElapsedTime(&a, s->t_ref, sl.t0) == hipSuccess &&
                hipEventElapsedTime(&b, s->t_ref, sl.t1) == hipSuccess && s->intervals.size() < (1u << 22))
                s->intervals.emplace_back(a, b);
        }
        for (Request& r : sl.reqs) finish(s, r, status);
        sl.reqs.clear();
        s->idle.push_back(si);
        s->busy.erase(s->busy.begin() + q);
        prune_done(s);
        any = true;
    }
    return any;
}

void worker_main(migrid_server* s)
{
    (void)hipSetDevice(s->cfg.device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        if (reap(s)) s->done_cv.notify_all();
        // requests that admission or batching fails finish here: their waiters
        // are woken below
        const int64_t finished = s->st.completed + s->st.failed;
        admit(s);
        bool launched = false;
        while (!s->admitted.empty() && !s->idle.empty()) {
            const int si = s->idle.back();
            std::vector<Request> take;
            const int rc = launch_batch(s, si, take);
            if (rc == MIGRID_OK && take.empty() && s->slots[si].reqs.empty()) break;  // all held by readers
            if (rc != MIGRID_OK) {
                snprintf(s->last_err, sizeof s->last_err, "%s", migrid_last_error());
                for (Request& r : take) finish(s, r, rc);
                s->done_cv.notify_all();
                continue;
            }
            s->idle.pop_back();
            s->busy.push_back(si);
            launched = true;
        }
        if (s->st.completed + s->st.failed != finished) s->done_cv.notify_all();
        if (launched) continue;
        if (s->stop && s->pending.empty() && s->admitted.empty() && s->busy.empty()) break;
        // poll the device queue; submissions wake the worker early
        s->cv.wait_for(lk, std::chrono::microseconds(s->busy.empty() ? 2000 : 50));
    }
}

// True when no request of the job is queued, admitted or running (mu held).
bool job_idle(const migrid_server* s, int64_t job_id)
{
    auto j = s->jobs.find(job_id);
    if (j != s->jobs.end() && j->second.inflight > 0) return false;
    for (const Request& r : s->pending)
        if (r.job_id == job_id) return false;
    return true;
}

}  // namespace

extern "C" {

void migrid_server_cfg_default(migrid_server_cfg* c)
{
    if (!c) return;
    memset(c, 0, sizeof *c);
    c->in_type = MIGRID_F64;
    c->out_type = MIGRID_F64;
    c->max_batch = 64;
    c->max_inflight = 2;
    migrid_opts_default(&c->opts);
    c->opts.strategy = MIGRID_TILED;
    c->opts.arith = MIGRID_ARITH_F32;
}

int migrid_server_create(const migrid_server_cfg* cfg, migrid_server** out)
{
    clear_error();
    if (!cfg || !out) return fail(MIGRID_EINVAL, "NULL argument");
    *out = nullptr;
    if ((cfg->in_type != MIGRID_F64 && cfg->in_type != MIGRID_F32) ||
        (cfg->out_type != MIGRID_F64 && cfg->out_type != MIGRID_F32))
        return fail(MIGRID_EINVAL, "server types %d/%d", cfg->in_type, cfg->out_type);
    if (cfg->max_batch < 0 || cfg->max_inflight < 0 || cfg->max_inflight > 8 || cfg->hbm_budget < 0)
        return fail(MIGRID_EINVAL, "max_batch=%d max_inflight=%d hbm_budget=%lld", cfg->max_batch,
                    cfg->max_inflight, (long long)cfg->hbm_budget);
    if (cfg->opts.debug_coverage) return fail(MIGRID_ENOTSUP, "the server does not run coverage mode");
    if (cfg->in_type == MIGRID_F64 && cfg->out_type == MIGRID_F32 &&
        cfg->opts.arith == MIGRID_ARITH_NATIVE)
        return fail(MIGRID_ENOTSUP, "double input into a float grid needs FP32 arithmetic");
    MG_HIP(hipSetDevice(cfg->device));
    migrid_server* s = new (std::nothrow) migrid_server;
    if (!s) return fail(MIGRID_ENOMEM, "host allocation of the server failed");
    s->cfg = *cfg;
    if (s->cfg.max_batch == 0) s->cfg.max_batch = 64;
    if (s->cfg.max_inflight == 0) s->cfg.max_inflight = 3;
    if (s->cfg.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete s;
            return fail(MIGRID_EDEVICE, "hipMemGetInfo failed");
        }
        s->cfg.hbm_budget = (int64_t)(fr / 10 * 9);
    }
    memset(&s->st, 0, sizeof s->st);
    // every stream at the default priority
    hipError_t e = hipStreamCreateWithFlags(&s->s_copy, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&s->t_ref);
    if (e == hipSuccess) e = hipEventRecord(s->t_ref, s->s_copy);
    if (e == hipSuccess) e = hipEventSynchronize(s->t_ref);
    s->slots.resize(s->cfg.max_inflight);
    for (int i = 0; i < s->cfg.max_inflight && e == hipSuccess; i++) {
        e = hipStreamCreateWithFlags(&s->slots[i].stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&s->slots[i].t0);
        if (e == hipSuccess) e = hipEventCreate(&s->slots[i].t1);
        s->idle.push_back(i);
    }
    if (e != hipSuccess) {
        const int rc = fail_hip(e, "server streams/events", __FILE__, __LINE__);
        for (Slot& sl : s->slots) {
            if (sl.t0) (void)hipEventDestroy(sl.t0);
            if (sl.t1) (void)hipEventDestroy(sl.t1);
            if (sl.stream) (void)hipStreamDestroy(sl.stream);
        }
        if (s->t_ref) (void)hipEventDestroy(s->t_ref);
        if (s->s_copy) (void)hipStreamDestroy(s->s_copy);
        delete s;
        return rc;
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return MIGRID_OK;
}

int migrid_server_submit(migrid_server* s, int64_t job_id, int64_t img_dim, int64_t gcf_dim,
                         const void* gcf, const void* uv, const void* vals, int64_t npts,
                         int32_t flags, int64_t* ticket)
{
    clear_error();
    if (!s || !ticket) return fail(MIGRID_EINVAL, "NULL argument");
    if (flags & ~(MIGRID_SUBMIT_DEVICE | MIGRID_SUBMIT_KEEP_SORT | MIGRID_SUBMIT_SAME_UV))
        return fail(MIGRID_EINVAL, "flags=%d", flags);
    MG_TRY(validate_shape(npts, img_dim, gcf_dim));
    if (gcf_dim > 64 || img_dim + gcf_dim - 2 >= 65536)
        return fail(MIGRID_ENOTSUP, "served jobs need S <= 64 and N + S - 2 < 65536");
    if (s->cfg.opts.gcf_mode != MIGRID_GCF_ANALYTIC && !gcf) return fail(MIGRID_EINVAL, "gcf is NULL");
    if (npts > 0 && (!uv || !vals)) return fail(MIGRID_EINVAL, "uv/vals is NULL");
    Request r;
    r.job_id = job_id;
    r.N = img_dim;
    r.S = gcf_dim;
    r.npts = npts;
    r.host = !(flags & MIGRID_SUBMIT_DEVICE);
    r.flags = flags;
    if (r.host) {
        // copied now into pinned staging: the caller may reuse its buffers
        const size_t e = s->cfg.in_type == MIGRID_F64 ? 16 : 8;
        MG_HIP(hipHostMalloc(&r.pinned, input_bytes(s->cfg, npts, gcf_dim), hipHostMallocDefault));
        if (npts > 0) {
            memcpy(r.pinned, uv, (size_t)npts * e);
            memcpy((char*)r.pinned + (size_t)npts * e, vals, (size_t)npts * e);
        }
        if (gcf)
            memcpy((char*)r.pinned + (size_t)npts * e * 2, gcf,
                   (size_t)kSubKernels * gcf_dim * gcf_dim * e);
    } else {
        r.uv = uv;
        r.vals = vals;
        r.gcf = gcf;
    }
    r.t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            if (r.pinned) (void)hipHostFree(r.pinned);
            return fail(MIGRID_EINVAL, "server is stopping");
        }
        r.ticket = s->next_ticket++;
        *ticket = r.ticket;
        s->st.submitted++;
        s->pending.push_back(r);
    }
    s->cv.notify_one();
    return MIGRID_OK;
}

int migrid_server_poll(migrid_server* s, int64_t ticket, int32_t* done, double* latency_ms)
{
    clear_error();
    if (!s || !done) return fail(MIGRID_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (ticket <= 0 || ticket >= s->next_ticket) return fail(MIGRID_EINVAL, "unknown ticket");
    if (ticket < s->released_below)
        return fail(MIGRID_EINVAL, "request %lld completed long ago; its status was released",
                    (long long)ticket);
    auto it = s->done.find(ticket);
    *done = it != s->done.end();
    if (!*done) return MIGRID_OK;
    if (latency_ms) *latency_ms = it->second.latency_ms;
    if (it->second.status != MIGRID_OK)
        return fail(it->second.status, "request %lld failed: %s", (long long)ticket, s->last_err);
    return MIGRID_OK;
}

int migrid_server_wait(migrid_server* s, int64_t ticket, double* latency_ms)
{
    clear_error();
    if (!s) return fail(MIGRID_EINVAL, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket <= 0 || ticket >= s->next_ticket) return fail(MIGRID_EINVAL, "unknown ticket");
    s->done_cv.wait(lk, [&] { return s->done.count(ticket) > 0 || ticket < s->released_below; });
    if (ticket < s->released_below)
        return fail(MIGRID_EINVAL, "request %lld completed long ago; its status was released",
                    (long long)ticket);
    const Done d = s->done[ticket];
    if (latency_ms) *latency_ms = d.latency_ms;
    if (d.status != MIGRID_OK)
        return fail(d.status, "request %lld failed: %s", (long long)ticket, s->last_err);
    return MIGRID_OK;
}

int migrid_server_read_grid(migrid_server* s, int64_t job_id, void* host_out)
{
    clear_error();
    if (!s || !host_out) return fail(MIGRID_EINVAL, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->done_cv.wait(lk, [&] { return job_idle(s, job_id); });
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end())
        return fail(MIGRID_EINVAL, "no resident grid for job %lld", (long long)job_id);
    // batches for this job are held back until the copy ends
    it->second.readers++;
    void* grid = it->second.grid;
    const size_t bytes = it->second.bytes;
    lk.unlock();
    const hipError_t e = hipMemcpy(host_out, grid, bytes, hipMemcpyDeviceToHost);
    lk.lock();
    s->jobs[job_id].readers--;
    lk.unlock();
    s->cv.notify_one();
    if (e != hipSuccess) return fail_hip(e, "read_grid copy", __FILE__, __LINE__);
    return MIGRID_OK;
}

int migrid_server_drop_job(migrid_server* s, int64_t job_id)
{
    clear_error();
    if (!s) return fail(MIGRID_EINVAL, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    s->done_cv.wait(lk, [&] {
        auto j = s->jobs.find(job_id);
        return job_idle(s, job_id) && (j == s->jobs.end() || j->second.readers == 0);
    });
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return MIGRID_OK;
    MG_HIP(hipFree(it->second.grid));
    s->hbm_used -= it->second.bytes;
    if (it->second.sort) {
        MG_HIP(hipFree(it->second.sort));
        s->hbm_used -= it->second.sort_bytes;
    }
    s->jobs.erase(it);
    s->cv.notify_one();
    return MIGRID_OK;
}

int migrid_server_get_stats(migrid_server* s, migrid_server_stats* out)
{
    clear_error();
    if (!s || !out) return fail(MIGRID_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->hbm_in_use = (int64_t)(s->hbm_used + arenas_bytes(s));
    out->hbm_budget = s->cfg.hbm_budget;
    out->jobs_resident = (int64_t)s->jobs.size();
    out->queued = (int64_t)(s->pending.size() + s->admitted.size());
    int64_t running = 0;
    for (int i : s->busy) running += (int64_t)s->slots[i].reqs.size();
    out->running = running;
    // union of the completed batches' device intervals, and the time two or
    // more of them overlapped
    std::vector<std::pair<double, double>> iv = s->intervals;
    std::sort(iv.begin(), iv.end());
    double busy = 0, overlap = 0, end = -1e300;
    for (const auto& x : iv) {
        if (x.first >= end) {
            busy += x.second - x.first;
            end = x.second;
        } else {
            overlap += std::min(x.second, end) - x.first;
            if (x.second > end) {
                busy += x.second - end;
                end = x.second;
            }
        }
    }
    out->gpu_busy_ms = busy;
    out->overlap_ms = overlap;
    return MIGRID_OK;
}

int migrid_server_destroy(migrid_server* s)
{
    clear_error();
    if (!s) return MIGRID_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();  // drains every submitted request
    for (auto& kv : s->jobs) {
        (void)hipFree(kv.second.grid);
        if (kv.second.sort) (void)hipFree(kv.second.sort);
    }
    for (Slot& sl : s->slots) {
        if (sl.arena) (void)hipFree(sl.arena);
        if (sl.bins) (void)hipFree(sl.bins);
        (void)hipEventDestroy(sl.t0);
        (void)hipEventDestroy(sl.t1);
        (void)hipStreamDestroy(sl.stream);
    }
    (void)hipEventDestroy(s->t_ref);
    (void)hipStreamDestroy(s->s_copy);
    delete s;
    return MIGRID_OK;
}

}  // extern "C"
