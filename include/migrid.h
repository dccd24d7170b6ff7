/*
 * migrid.h -- C ABI of the MI355X (gfx950) convolutional gridder.
 *
 * Drop-in replacement for the reference's gridding boundary:
 *   template <class CmplxOutType, class CmplxType>
 *   void gridGPU(CmplxOutType* out, CmplxType* in, CmplxType* in_vals,
 *                size_t npts, size_t img_dim, CmplxType* gcf, size_t gcf_dim);
 *        (grid_gpu.cuh:30-32, body grid_gpu.cu:675-819, instantiated for
 *         <double2,double2>, <double2,float2>, <float2,float2> at :820-825)
 *   void gridCPU   (PRECISION2* out, ...)      (grid.cu:87-139)
 *   void gridCPU_pz(OUTPRECISION2* out, ...)   (grid.cu:140-193)
 *
 * Plain C types only: complex values are {x, y} pairs laid out like
 * float2/double2.  Every entry point returns 0 on success or a negative
 * MIGRID_E* code; migrid_last_error() gives the message (per thread).
 * Oversampling is fixed at 8 (64 sub-kernels), one polarisation, as in the
 * reference (Defines.h:60,77).  img_dim and gcf_dim are run-time values.
 */
#ifndef MIGRID_H
#define MIGRID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y; } migrid_c128; /* layout of double2 */
typedef struct { float x, y; } migrid_c64;   /* layout of float2 */

/* ---- status codes ------------------------------------------------------- */
#define MIGRID_OK 0
#define MIGRID_EINVAL (-1)  /* bad argument or shape */
#define MIGRID_ENOMEM (-2)  /* device or host allocation failed */
#define MIGRID_EDEVICE (-3) /* HIP runtime error (message has the HIP string) */
#define MIGRID_ENOTSUP (-4) /* combination not supported */
#define MIGRID_EBUSY (-5)   /* admission refused: not enough free HBM */

/* ---- options ------------------------------------------------------------ */
/* Strategy.  SCATTER is the reference default (grid_gpu.cu:778-786): one
 * hardware atomic per tap.  GATHER and WINDOW are the reference's other two
 * (:734-776), rebuilt as per-tile owner kernels that use no atomics: a dense
 * tile split over several work items is summed from their partial tiles in
 * a fixed order, so the grid is the same bit for bit on every run (for one
 * grid per job; jobs of a batch sharing a grid flush atomically).  TILED is
 * the bin-sorted LDS-tiled gridder with atomics only where a split tile is
 * flushed; at supports up to opts.tap_smax (16 by default) it sums each tile
 * one tap per lane with LDS atomics instead (the tap kernel), 2-5.7x faster
 * there. */
#define MIGRID_SCATTER 0
#define MIGRID_GATHER 1
#define MIGRID_WINDOW 2
#define MIGRID_TILED 3

/* Output layout.  DEFAULT follows the reference per strategy: interleaved
 * {re,im} for scatter/gather (grid_gpu.cu:162-163,500), planar (real plane
 * then imaginary plane) for the moving window (:524-525,585-586) and for the
 * gridCPU entry points (grid.cu:101-102,153-154).  TILED defaults to
 * interleaved. */
#define MIGRID_LAYOUT_DEFAULT 0
#define MIGRID_LAYOUT_INTERLEAVED 1
#define MIGRID_LAYOUT_PLANAR 2

/* GCF source.  TABLE reads gcf[64*S*S].  COMPAT is the reference's
 * COMPUTE_GCF (grid_gpu.cu:103-108,137-147): T = gcf[S(S-1)/2-1].x,
 * w = .y, every tap of a visibility uses (sinf, cosf)(phase).  ANALYTIC
 * evaluates exp(-(dx^2+dy^2)/(2 sigma^2)) * exp(i beta (dx^2+dy^2)) at the
 * table's quantised offsets on the device (no table is read). */
#define MIGRID_GCF_TABLE 0
#define MIGRID_GCF_COMPAT 1
#define MIGRID_GCF_ANALYTIC 2

/* Tile arithmetic (TILED strategy).  NATIVE: products in the input type,
 * sums in the output type.  F32: vis x GCF products and in-tile sums in
 * FP32.  F32_BF16GCF: the same with the GCF staged as bfloat16. */
#define MIGRID_ARITH_NATIVE 0
#define MIGRID_ARITH_F32 1
#define MIGRID_ARITH_F32_BF16GCF 2

typedef struct migrid_opts {
    int32_t strategy;       /* MIGRID_SCATTER .. MIGRID_TILED */
    int32_t layout;         /* MIGRID_LAYOUT_* */
    int32_t gcf_mode;       /* MIGRID_GCF_* */
    int32_t arith;          /* MIGRID_ARITH_* (TILED only) */
    int32_t accumulate;     /* 0: fresh grid (zeroed first), 1: add into it */
    int32_t debug_coverage; /* 1: every tap adds (1, n) -- the reference's
                               DEBUG1 coverage map (grid.cu:127-129) */
    int32_t tile;           /* TILED/GATHER/WINDOW tile: 16 or 32 (square tiles of
                               16 x 16 wave blocks), 64 (128 x 64 tiles of 8 x 64
                               wave blocks), 0 = auto (64) */
    int32_t vis_per_wave;   /* TILED work-item size, 0 = auto */
    double gcf_sigma;       /* ANALYTIC: Gaussian sigma in pixels (0: S/6) */
    double gcf_beta;        /* ANALYTIC: quadratic phase coefficient */
    int32_t stage_min;      /* TILED tuning, square tiles (tile 16 or 32): a
                               sub-kernel with fewer visibilities for a tile is
                               read from HBM instead of staged in LDS (0 = auto);
                               the 128 x 64 tiles stage every sub-kernel */
    int32_t wg_per_cu;      /* TILED tuning: persistent workgroups per CU (0 = auto) */
    int32_t tap_smax;       /* TILED: supports up to this (at most 16) are gridded
                               one tap per lane into an LDS tile; 0 = 16, -1 = never */
    int32_t reserved[5];
} migrid_opts;

/* Fill *o with the defaults (scatter, default layout, table GCF, fresh grid). */
void migrid_opts_default(migrid_opts* o);

/* Message for the last failing call on this thread ("" if none). */
const char* migrid_last_error(void);
const char* migrid_version(void);

/* ---- drop-in host entry points ------------------------------------------ */
/* gridGPU<double2,double2> / <double2,float2> / <float2,float2>
 * (grid_gpu.cu:675-819).  `out` is the PADDED host buffer of
 * (img_dim^2 + 2*img_dim*gcf_dim + 2*gcf_dim) complex elements; the grid base
 * is out + img_dim*gcf_dim + gcf_dim (grid_gpu.cu:711,732).  The whole padded
 * buffer is zeroed on the device, gridded and copied back (unless
 * opts->accumulate, which grids into the values already in `out`).
 * in[n] = {u, v} in pixels, in_vals[n] paired by index, gcf[64*gcf_dim^2].
 * opts may be NULL (reference semantics: scatter, interleaved). */
int migrid_gridGPU_zz(migrid_c128* out, const migrid_c128* in, const migrid_c128* in_vals,
                      size_t npts, size_t img_dim, const migrid_c128* gcf, size_t gcf_dim,
                      const migrid_opts* opts);
int migrid_gridGPU_zc(migrid_c128* out, const migrid_c64* in, const migrid_c64* in_vals,
                      size_t npts, size_t img_dim, const migrid_c64* gcf, size_t gcf_dim,
                      const migrid_opts* opts);
int migrid_gridGPU_cc(migrid_c64* out, const migrid_c64* in, const migrid_c64* in_vals,
                      size_t npts, size_t img_dim, const migrid_c64* gcf, size_t gcf_dim,
                      const migrid_opts* opts);

/* gridCPU (grid.cu:87-139): `out` is the UNPADDED grid base; the result is
 * PLANAR (real plane of img_dim^2 doubles, then the imaginary plane), zeroed
 * first.  Computed on the GPU.  gridCPU_pz (grid.cu:140-193) is the same but
 * adds into the planes already in `out`.  _zz: double everything; _zc:
 * float inputs, double planes; _cc: float inputs, float planes. */
int migrid_gridCPU_zz(migrid_c128* out, const migrid_c128* in, const migrid_c128* in_vals,
                      size_t npts, size_t img_dim, const migrid_c128* gcf, size_t gcf_dim);
int migrid_gridCPU_pz_zz(migrid_c128* out, const migrid_c128* in, const migrid_c128* in_vals,
                         size_t npts, size_t img_dim, const migrid_c128* gcf, size_t gcf_dim);
int migrid_gridCPU_pz_zc(migrid_c128* out, const migrid_c64* in, const migrid_c64* in_vals,
                         size_t npts, size_t img_dim, const migrid_c64* gcf, size_t gcf_dim);
int migrid_gridCPU_pz_cc(migrid_c64* out, const migrid_c64* in, const migrid_c64* in_vals,
                         size_t npts, size_t img_dim, const migrid_c64* gcf, size_t gcf_dim);

/* Size in complex elements of the padded buffer gridGPU expects. */
size_t migrid_padded_elems(size_t img_dim, size_t gcf_dim);

/* ---- device-resident entry points --------------------------------------- */
/* Element types of a device job. */
#define MIGRID_F64 0
#define MIGRID_F32 1

/* One gridding job on device memory.  All pointers are device pointers.
 * grid is the UNPADDED base: N*N complex interleaved, or two planes of N*N
 * (real then imaginary) when the layout is planar.  grid_lo is the
 * compensation grid of the FP32 two-sum flush (same shape as grid), or NULL. */
typedef struct migrid_job {
    int32_t in_type;  /* MIGRID_F64 / MIGRID_F32: uv, vals and gcf */
    int32_t out_type; /* MIGRID_F64 / MIGRID_F32: grid */
    const void* uv;   /* npts x {u,v} */
    const void* vals; /* npts x {re,im} */
    const void* gcf;  /* 64*S*S x {re,im} (may be NULL for ANALYTIC) */
    void* grid;
    void* grid_lo;
    int64_t npts;
    int64_t img_dim;
    int64_t gcf_dim;
} migrid_job;

/* A context owns device scratch (bin-sort buffers, work queues), reused
 * across calls.  One context is used by one host thread at a time; calls on
 * one context are ordered on the stream passed to them. */
typedef struct migrid_ctx migrid_ctx;

int migrid_ctx_create(migrid_ctx** ctx, int device);
int migrid_ctx_destroy(migrid_ctx* ctx);

/* Grid one job on `stream` (a hipStream_t, or NULL for the null stream).
 * Asynchronous with respect to the host. */
int migrid_ctx_grid(migrid_ctx* ctx, const migrid_job* job, const migrid_opts* opts,
                    void* stream);

/* Grid one job like migrid_ctx_grid, allocating nothing.  No counterpart
 * in the reference: this is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The tiled strategies (TILED, GATHER, WINDOW) take their device
 * scratch -- bin sort, records, work queue, partial tiles -- from one static
 * device array of MIGRID_NOALLOC_SCRATCH_BYTES per GPU, which the library
 * holds until the process ends; a job that needs more fails with
 * MIGRID_ENOTSUP before any launch (grid it through a context instead).
 * SCATTER needs no scratch.  All calls on one GPU share that array: calls
 * that overlap on one GPU race, so callers serialise them.  No context, no
 * profiling events.  Asynchronous on stream (a hipStream_t, or NULL). */
#define MIGRID_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int migrid_grid_noalloc(const migrid_job* job, const migrid_opts* opts, void* stream);

/* Bin sort only (device pointers): stable order of visibility indices by
 * key (k, bin_y, bin_x) with k = 8*sub_y + sub_x, bins of `bin` pixels over
 * q = floor(u) + S/2 in [0, N+S-2]; perm[0..count) and CSR offsets
 * [0..64*nb*nb] with nb = (N+S-2)/bin + 1.  *count receives the number of
 * binned visibilities (those touching the grid).  Synchronous. */
int migrid_ctx_bin_sort(migrid_ctx* ctx, int32_t in_type, const void* uv, int64_t npts,
                        int64_t img_dim, int64_t gcf_dim, int32_t bin, int32_t* perm,
                        int32_t* offsets, int64_t* count, void* stream);

/* Kernel timing: when enabled, every call records HIP events (on the call's
 * stream) around the main gridding kernel.  migrid_ctx_profile_read
 * synchronises on them and copies up to `cap` per-launch durations (ms) into
 * ms_out (may be NULL); it returns the number of launches recorded since the
 * last reset, or a negative status. */
int migrid_ctx_set_profiling(migrid_ctx* ctx, int enable);
int migrid_ctx_profile_read(migrid_ctx* ctx, double* ms_out, int cap);
int migrid_ctx_profile_reset(migrid_ctx* ctx);

/* ---- batched launch (many jobs, one persistent kernel) --------------------
 * No counterpart in the reference, which grids one job per gridGPU call.
 * A batch plans all jobs (which must share input/output types, layout, GCF
 * mode and tile edge; sizes, supports and grids may differ) into one device
 * arena at creation.  migrid_batch_run enqueues, on `stream`: the grid clears
 * (unless opts->accumulate), every job's bin sort, record packing and work
 * items, then ONE persistent owner-kernel launch over a shared device work
 * queue (and, for TILED batches with supports up to opts.tap_smax, one
 * tap-kernel launch over those jobs' own queue).  migrid_batch_graph_launch
 * captures that same sequence into a
 * hipGraph on its first call (stream must not be NULL) and replays it after;
 * new data may be written into the same job buffers between replays.  The
 * job pointers given at creation must stay valid for the batch's lifetime. */
typedef struct migrid_batch migrid_batch;

int migrid_batch_create(migrid_ctx* ctx, const migrid_job* jobs, int njobs,
                        const migrid_opts* opts, migrid_batch** batch);
int migrid_batch_run(migrid_batch* batch, void* stream);
int migrid_batch_graph_launch(migrid_batch* batch, void* stream);
int migrid_batch_info(migrid_batch* batch, int64_t* arena_bytes, int32_t* tile, int32_t* njobs);
int migrid_batch_destroy(migrid_batch* batch);

/* Autotuner: times tile edge {16, 32, 64} x visibilities per wavefront
 * {64, 128, 256} (a work item holds 64 x that many visibilities) on the given
 * jobs, then the LDS split around the fastest: 1 or 2 workgroups per CU
 * (opts.wg_per_cu) and the staging threshold (opts.stage_min 2 or 16, against
 * the default 4); returns the fastest options in *best (and its time in *best_ms), and
 * caches the choice per shape class (set of supports and grid sizes,
 * arithmetic, types, visibility-count decade) for the life of the process.
 * Synchronous. */
int migrid_autotune(migrid_ctx* ctx, const migrid_job* jobs, int njobs, const migrid_opts* base,
                    migrid_opts* best, double* best_ms, void* stream);

/* ---- measured ceilings ---------------------------------------------------
 * No counterpart in the reference, which prints only its own kernels'
 * rates (grid_gpu.cu:789-791).  Synchronous, on the current device.
 * migrid_ceiling_lds: LDS read bandwidth at the tiled owner kernel's
 * occupancy (1024-thread workgroups, two per CU), elem_bytes 8 (FP32 pairs)
 * or 16 (FP64 pairs); pitch 0: linear rows (lane i reads element i), else
 * the owner kernel's lane block (2 x 2 elements at stride 8) with that row
 * pitch in elements (96..124 in steps of 4, 97 or 99), shifted per step;
 * amount = bytes read; per_clk_cu and clock_ghz count s_memtime ticks.
 * migrid_ceiling_atomics: the plain scatter kernel's global atomics alone
 * (2 per tap, the job's visibilities and grid; constants added, no table
 * reads): amount = atomics issued, assuming every tap on the grid. */
typedef struct migrid_ceiling {
    double ms;         /* kernel time of the measured launch (HIP events) */
    double amount;     /* bytes read (LDS) or atomics issued */
    double per_s;      /* amount per second */
    double per_clk_cu; /* LDS: bytes per s_memtime tick per CU */
    double clock_ghz;  /* LDS: s_memtime ticks per ns over the loop */
} migrid_ceiling;
int migrid_ceiling_lds(int32_t elem_bytes, int32_t pitch, int32_t iters, migrid_ceiling* res);
int migrid_ceiling_atomics(const migrid_job* job, migrid_ceiling* res);

/* Number of bins per axis for migrid_ctx_bin_sort. */
int64_t migrid_bins_per_axis(int64_t img_dim, int64_t gcf_dim, int32_t bin);

/* ---- input side: coordinates and per-visibility GCF index ------------------
 * (u, v, w) in metres per row and channel frequencies in Hz -> grid pixel
 * coordinates for the gridder, visibility n = row * nchan + chan (the order
 * of the reference's HDF5 reader, hdf52struct.cpp:95-113):
 *   u_pix = img_dim/2 + (u f / c) img_dim cell_rad, v_pix likewise,
 * c = 299792458 m/s, cell_rad the image pixel size in radians.  With
 * nplanes > 0 also a w-plane GCF index per visibility:
 *   g = min(nplanes - 1, floor(|w f / c| / w_max * nplanes)).
 * The reference's device convert() (grid_gpu.cu:45-51) is the normalised
 * special case; its vis.gcfinx (vis.h:17) is an index of this kind, which no
 * reference kernel reads.  Device pointers (uvw: nrow x 3 doubles, freq:
 * nchan doubles, uv_out: nrow*nchan x 2 doubles); asynchronous on stream. */
int migrid_uvw_to_pixels(const double* uvw, int64_t nrow, const double* freq, int32_t nchan,
                         int64_t img_dim, double cell_rad, int32_t nplanes, double w_max,
                         double* uv_out, int32_t* gcf_index_out, void* stream);

/* Gridding with one GCF table per visibility: job->gcf holds ngcf tables of
 * 64 * gcf_dim^2 complex values one after another, and gcf_index[n] in
 * [0, ngcf) picks visibility n's table (device int32 array).  The
 * visibilities are grouped by table on the device and gridded as one
 * persistent launch of the owner kernel, or of the tap kernel at supports up
 * to opts.tap_smax under TILED (strategy TILED, GATHER or WINDOW; SCATTER is
 * taken as TILED), all groups flushing atomically into the one grid.  GCF mode TABLE only.  Synchronises with stream once (to read the
 * group sizes); the gridding itself is asynchronous. */
int migrid_ctx_grid_indexed(migrid_ctx* ctx, const migrid_job* job, int32_t ngcf,
                            const int32_t* gcf_index, const migrid_opts* opts, void* stream);

/* Degridding, the adjoint of gridding (no counterpart in the reference):
 *   vals_out[n] = sum over the in-grid taps of visibility n of
 *                 conj(G[k][r][c]) * grid[main_y - S/2 + 1 + r][main_x - S/2 + 1 + c]
 * with the gridder's index math, so <grid(v), g> = <v, degrid(g)>.  job->grid
 * is read (N x N complex, interleaved unless opts->layout is PLANAR),
 * job->vals is ignored and vals_out (npts complex of in_type) is written;
 * visibilities whose window misses the grid get 0.  GCF mode TABLE; arith
 * NATIVE (double products for double visibilities) or F32.  S <= 64.
 * Asynchronous on stream. */
int migrid_ctx_degrid(migrid_ctx* ctx, const migrid_job* job, void* vals_out,
                      const migrid_opts* opts, void* stream);

/* Grid <-> image: the centred, unnormalised 2-D FFT (no counterpart in the
 * reference, whose path ends at the grid).  direction MIGRID_FFT_TO_IMAGE:
 *   img[y][x] = scale / (taper[x] taper[y]) *
 *     sum_{v,u} grid[v][u] exp(+2 pi i ((u - N/2)(x - N/2) + (v - N/2)(y - N/2)) / N)
 * i.e. numpy's fftshift(ifft2(ifftshift(grid))) * N^2 * scale / taper;
 * MIGRID_FFT_TO_GRID is its adjoint (the input divided by the taper, then the
 * exp(-...) transform, times scale).  taper: NULL or N device doubles (the
 * separable grid correction).  in: N x N complex of `type` (interleaved, or
 * planar (re plane, im plane) with layout PLANAR); out: N x N interleaved
 * complex, may equal an interleaved `in`.  N even.  One hipFFT plan is cached
 * per context.  Asynchronous on stream. */
#define MIGRID_FFT_TO_IMAGE 1
#define MIGRID_FFT_TO_GRID (-1)
int migrid_ctx_fft(migrid_ctx* ctx, int32_t type, int32_t layout, const void* in, void* out,
                   int64_t img_dim, int32_t direction, double scale, const double* taper,
                   void* stream);

/* The reference's UVW text input (grid.cu:338-366): lines
 * "%d,%d,%d: %f, %f, %f" (three integers, u, v, w), u and v scaled to the
 * grid as the reference does, x = (double)(float)(u * (float)img_dim) / 2048.
 * Host memory: reads up to max_rows lines into uv_out (max_rows x 2) and,
 * when w_out is not NULL, w into w_out; *nrows_out = the rows read.  Blank
 * lines are skipped; any other unreadable line is an error (EINVAL). */
int migrid_read_uvw_text(const char* path, int64_t img_dim, int64_t max_rows, double* uv_out,
                         double* w_out, int64_t* nrows_out);

/* ---- serving loop (continuous batching) -----------------------------------
 * No counterpart in the reference, whose only entry is the blocking
 * per-job gridGPU call (grid_gpu.cu:675-819).  A server owns a worker thread
 * on one device, a copy stream and one stream per batch slot.  Requests carry
 * a job id; each job id owns a resident N x N grid (interleaved complex,
 * out_type) that starts at zero and accumulates every request submitted for
 * it until migrid_server_drop_job.  Requests are admitted in arrival order
 * while their device footprint fits hbm_budget, and everything admitted forms
 * the next batch (up to max_batch requests, one per job id) as soon as a slot
 * is free: up to max_inflight batches run on the device at once, each one
 * persistent launch of the tiled gridder with cfg.opts (strategy and
 * accumulate are forced) on its own stream.  Batches that write the same grid
 * run in submission order; others overlap.  A job can keep its bin sort
 * resident (MIGRID_SUBMIT_KEEP_SORT) so later requests with the same
 * coordinates (MIGRID_SUBMIT_SAME_UV) skip the sort and only pack their new
 * values.  Thread-safe: any thread may submit, poll, wait, read and drop. */
typedef struct migrid_server migrid_server;

typedef struct migrid_server_cfg {
    int32_t device;
    int32_t in_type;      /* MIGRID_F64 or MIGRID_F32 uv/vals/gcf */
    int32_t out_type;     /* resident grid element type */
    int32_t max_batch;    /* requests per launch (0: 64) */
    int32_t max_inflight; /* batches on the device at once, 1..8 (0: 3) */
    int32_t reserved0;
    int64_t hbm_budget;   /* bytes the server may hold (0: 90% of free HBM) */
    migrid_opts opts;     /* arithmetic, tile, tuning; GCF mode */
} migrid_server_cfg;

typedef struct migrid_server_stats {
    int64_t submitted, completed, failed, batches;
    int64_t vis_completed;
    int64_t hbm_in_use, hbm_budget, jobs_resident;
    int64_t queued;  /* submitted or admitted, not yet on the device */
    int64_t running; /* in batches on the device */
    double gpu_ms;   /* summed device time of completed batches */
    double gpu_busy_ms;  /* time with at least one batch on the device (union) */
    double overlap_ms;   /* time with two or more batches on the device */
    int64_t sorts_reused; /* requests that took their job's kept sort */
} migrid_server_stats;

#define MIGRID_SUBMIT_HOST 0   /* uv/vals/gcf in host memory, copied at submit */
#define MIGRID_SUBMIT_DEVICE 1 /* device pointers, valid until the request completes */
/* Keep this request's bin sort (its permutation and CSR offsets, about
 * 4 * npts + 4 * 64 * ((N + S - 2) / 16 + 1)^2 bytes) resident with its job,
 * replacing a sort kept before. */
#define MIGRID_SUBMIT_KEEP_SORT 2
/* This request's coordinates are those of the sort its job keeps (the same
 * count, support and values: the caller's promise): the sort is reused and
 * only the values are packed.  The request fails (EINVAL) when its job keeps
 * no sort of that count and support. */
#define MIGRID_SUBMIT_SAME_UV 4

void migrid_server_cfg_default(migrid_server_cfg* cfg);
int migrid_server_create(const migrid_server_cfg* cfg, migrid_server** server);
/* Queues one request; *ticket identifies it.  gcf holds 64 x gcf_dim^2
 * complex values as in gridGPU; a job id keeps its img_dim for life. */
int migrid_server_submit(migrid_server* server, int64_t job_id, int64_t img_dim, int64_t gcf_dim,
                         const void* gcf, const void* uv, const void* vals, int64_t npts,
                         int32_t flags, int64_t* ticket);
/* *done = 1 when the request has completed; returns its status then.
 * latency_ms: submit to completion as seen by the worker.  Statuses are kept
 * for repeated reads; past 65536 completed requests the oldest (all below
 * every pending ticket) are released down to half that, and poll/wait on a
 * released ticket return MIGRID_EINVAL ("released") instead of blocking. */
int migrid_server_poll(migrid_server* server, int64_t ticket, int32_t* done, double* latency_ms);
int migrid_server_wait(migrid_server* server, int64_t ticket, double* latency_ms);
/* Waits for every request of the job submitted so far, then copies its grid
 * (img_dim^2 complex of out_type) to host memory. */
int migrid_server_read_grid(migrid_server* server, int64_t job_id, void* host_out);
/* Waits like read_grid, then frees the job's grid. */
int migrid_server_drop_job(migrid_server* server, int64_t job_id);
int migrid_server_get_stats(migrid_server* server, migrid_server_stats* stats);
/* Completes every submitted request, then frees everything. */
int migrid_server_destroy(migrid_server* server);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* MIGRID_H */
