/*
 * oracle/gridcpu.c -- CPU restatement of the reference gridder's oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * as the timed CPU baseline.  The product (libmigrid.so) never links or calls
 * anything in this directory.
 *
 * What is restated, with the reference lines each function follows:
 *   oracle_grid_{zz,zc,cc}   gridCPU / gridCPU_pz        grid.cu:87-139, :140-193
 *   oracle_grid_zz_rows      the same loop restricted to a band of output rows
 *                            (per-pixel summation order unchanged, so the
 *                            OpenMP row-band baseline is bit-identical)
 *   oracle_grid_zz_omp,      row bands on OpenMP threads, rows outside and
 *   oracle_grid_const_tap_omp  columns inside; per-pixel order unchanged
 *   oracle_init_gcf          init_gcf                     grid.cu:72-85
 *   oracle_ref_inputs        default input generator      grid.cu:444-452
 *   oracle_vis2ints          vis2ints                     grid_gpu.cu:352-365
 *   oracle_set_bookmarks     set_bookmarks (without the   grid_gpu.cu:366-390
 *                            out-of-bounds read and overfill)
 *   oracle_bin_sort          the build's (k, bin_y, bin_x) stable counting
 *                            sort; it generalises set_bookmarks' cell CSR
 *   oracle_window_pixel      moving-window owner pixel    grid_gpu.cu:568-573
 *   oracle_compute_gcf_phase COMPUTE_GCF phase            grid_gpu.cu:103-108,137-147
 * plus the synthetic generators (SplitMix64 + Box-Muller) and the
 * Gaussian x quadratic-phase GCF table used by the benchmark workloads.
 *
 * Layout conventions (grid.cu:101-102,153-154): the CPU oracle writes a PLANAR
 * grid, a real plane of N*N values followed by an imaginary plane of N*N.
 * Taps: a and b run from S/2 down to -S/2+1 (b innermost); tap (a,b) of
 * sub-kernel k = 8*sub_y + sub_x reads gcf[k*S*S + (b+S/2-1)*S + (a+S/2-1)]
 * (grid.cu:100,115-120).  Built with -ffp-contract=off so no FMA contraction
 * changes the rounding of r1*r2-i1*i2.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OVERSAMPLE 8 /* GCF_GRID, Defines.h:60 */

/* ---- index math ----------------------------------------------------------
 * GPU form (grid_gpu.cu:118-121): exact double floor.
 * gridCPU form (grid.cu:107-111): sub goes through floorf on float-rounded
 * values.  The two agree whenever u is representable in float32.          */
static inline void index_exact(double u, int* main_, int* sub)
{
    double f = floor(u);
    *main_ = (int)f;
    *sub = (int)floor(OVERSAMPLE * (u - f));
}

static inline void index_floorf(double u, int* main_, int* sub)
{
    *sub = (int)floorf((float)(OVERSAMPLE * (u - (double)floorf((float)u))));
    *main_ = (int)floor(u);
}

/* ---- gridCPU / gridCPU_pz, double products, double planar grid ---------- */
static void grid_zz_band(double* out, const double* uv, const double* vals,
                         int64_t npts, int64_t N, const double* gcf, int64_t S,
                         int floorf_emul, int debug1, int64_t y_lo, int64_t y_hi)
{
    double* re = out;
    double* im = out + N * N;
    const int64_t h = S / 2;
    for (int64_t n = 0; n < npts; n++) {
        int mx, my, sx, sy;
        if (floorf_emul) {
            index_floorf(uv[2 * n], &mx, &sx);
            index_floorf(uv[2 * n + 1], &my, &sy);
        } else {
            index_exact(uv[2 * n], &mx, &sx);
            index_exact(uv[2 * n + 1], &my, &sy);
        }
        /* k may be -1 or 64 under floorf emulation; reading outside the table
         * would be undefined, so such a visibility is skipped here and the
         * caller is told through the return of oracle_floorf_divergent(). */
        int64_t k = (int64_t)OVERSAMPLE * sy + sx;
        if (k < 0 || k >= OVERSAMPLE * OVERSAMPLE) continue;
        const double* g = gcf + 2 * k * S * S;
        const double r1 = vals[2 * n], i1 = vals[2 * n + 1];
        for (int64_t a = h; a > -h; a--) {
            const int64_t x = mx + a;
            if (x < 0 || x >= N) continue;
            for (int64_t b = h; b > -h; b--) {
                const int64_t y = my + b;
                if (y < y_lo || y >= y_hi || y < 0 || y >= N) continue;
                const int64_t t = (b + h - 1) * S + (a + h - 1);
                const double r2 = g[2 * t], i2 = g[2 * t + 1];
                if (debug1) {
                    re[x + N * y] += 1.0;
                    im[x + N * y] += (double)n;
                } else {
                    re[x + N * y] += r1 * r2 - i1 * i2;
                    im[x + N * y] += r1 * i2 + r2 * i1;
                }
            }
        }
    }
}

/* zero_first=1 is gridCPU (zeroes N*N complex first, grid.cu:98);
 * zero_first=0 is gridCPU_pz (accumulates into the caller's grid). */
void oracle_grid_zz(double* out, const double* uv, const double* vals, int64_t npts,
                    int64_t N, const double* gcf, int64_t S, int floorf_emul,
                    int zero_first, int debug1)
{
    if (zero_first) memset(out, 0, sizeof(double) * 2 * N * N);
    grid_zz_band(out, uv, vals, npts, N, gcf, S, floorf_emul, debug1, 0, N);
}

/* Row band [y_lo, y_hi) only; used by the multi-threaded CPU baseline. */
void oracle_grid_zz_rows(double* out, const double* uv, const double* vals,
                         int64_t npts, int64_t N, const double* gcf, int64_t S,
                         int64_t y_lo, int64_t y_hi)
{
    grid_zz_band(out, uv, vals, npts, N, gcf, S, 0, 0, y_lo, y_hi);
}

/* Rows [y_lo, y_hi) from the listed visibilities only (ascending n), each
 * visibility visiting just its rows inside the band.  A visibility adds at
 * most once to a pixel, so the order of its taps does not matter: rows are
 * walked outside and columns inside (contiguous in each plane), and every
 * pixel still receives its contributions in ascending n, bit-identical to
 * the serial loop of grid.cu:115-133.  r2 == NULL: table taps; otherwise the
 * constant tap (r2[n], i2[n]) of COMPUTE_GCF, as oracle_grid_const_tap. */
__attribute__((optimize("tree-vectorize"))) static void grid_zz_list(double* out, const double* uv, const double* vals, const int64_t* list,
                         int64_t count, int64_t N, const double* gcf, int64_t S, const float* r2c,
                         const float* i2c, int64_t y_lo, int64_t y_hi)
{
    double* re = out;
    double* im = out + N * N;
    const int64_t h = S / 2;
    for (int64_t i = 0; i < count; i++) {
        const int64_t n = list[i];
        int mx, my, sx, sy;
        index_exact(uv[2 * n], &mx, &sx);
        index_exact(uv[2 * n + 1], &my, &sy);
        const int64_t k = (int64_t)OVERSAMPLE * sy + sx;
        const double* g = gcf ? gcf + 2 * k * S * S : NULL;
        const double r1 = vals[2 * n], i1 = vals[2 * n + 1];
        int64_t b_hi = y_hi - 1 - my, b_lo = y_lo - my; /* rows of the band */
        if (b_hi > h) b_hi = h;
        if (b_lo < -h + 1) b_lo = -h + 1;
        int64_t a_lo = -mx, a_hi = N - 1 - mx; /* columns on the grid */
        if (a_lo < -h + 1) a_lo = -h + 1;
        if (a_hi > h) a_hi = h;
        if (r2c) {
            const double pr = r1 * r2c[n] - i1 * i2c[n];
            const double pi = r1 * i2c[n] + r2c[n] * i1;
            for (int64_t b = b_lo; b <= b_hi; b++) {
                double* rr = re + N * (my + b) + mx;
                double* ri = im + N * (my + b) + mx;
                for (int64_t a = a_lo; a <= a_hi; a++) {
                    rr[a] += pr;
                    ri[a] += pi;
                }
            }
            continue;
        }
        for (int64_t b = b_lo; b <= b_hi; b++) {
            double* rr = re + N * (my + b) + mx;
            double* ri = im + N * (my + b) + mx;
            const double* gt = g + 2 * ((b + h - 1) * S + (h - 1));
            for (int64_t a = a_lo; a <= a_hi; a++) {
                const double gr = gt[2 * a], gi = gt[2 * a + 1];
                rr[a] += r1 * gr - i1 * gi;
                ri[a] += r1 * gi + gr * i1;
            }
        }
    }
}

/* Multi-threaded form: OpenMP over bands of 16 rows.  One serial pass lists
 * each band's visibilities in ascending n; each thread then grids its bands
 * from their lists.  Every pixel still receives its contributions in
 * ascending n, so the result is bit-identical to the serial loop.  Returns
 * the number of threads used (0 if the lists could not be allocated). */
static int grid_omp(double* out, const double* uv, const double* vals, int64_t npts, int64_t N,
                    const double* gcf, int64_t S, const float* r2c, const float* i2c, int nthreads)
{
    memset(out, 0, sizeof(double) * 2 * N * N);
    const int64_t band = 16, nb = (N + band - 1) / band, h = S / 2;
    int64_t* start = (int64_t*)calloc((size_t)nb + 1, sizeof(int64_t));
    if (!start) return 0;
    /* the band range of visibility n's rows [my - h + 1, my + h] on the grid;
     * empty when its window misses the grid's columns */
#define OMP_BANDS(n, b0, b1)                                                     \
    int mx_, my_, sx_, sy_;                                                      \
    index_exact(uv[2 * (n)], &mx_, &sx_);                                        \
    index_exact(uv[2 * (n) + 1], &my_, &sy_);                                    \
    int64_t r0_ = (int64_t)my_ - h + 1, r1_ = (int64_t)my_ + h;                  \
    if (r0_ < 0) r0_ = 0;                                                        \
    if (r1_ > N - 1) r1_ = N - 1;                                                \
    const int cols_ = (int64_t)mx_ + h >= 0 && (int64_t)mx_ - h + 1 <= N - 1;    \
    const int64_t k_ = (int64_t)OVERSAMPLE * sy_ + sx_;                          \
    const int64_t b0 = r0_ / band,                                               \
                  b1 = (r0_ <= r1_ && cols_ && k_ >= 0 && k_ < 64) ? r1_ / band : b0 - 1
    for (int64_t n = 0; n < npts; n++) {
        OMP_BANDS(n, b0, b1);
        for (int64_t b = b0; b <= b1; b++) start[b + 1]++;
    }
    for (int64_t b = 0; b < nb; b++) start[b + 1] += start[b];
    int64_t* list = (int64_t*)malloc((size_t)(start[nb] > 0 ? start[nb] : 1) * sizeof(int64_t));
    int64_t* fill = (int64_t*)malloc((size_t)nb * sizeof(int64_t));
    if (!list || !fill) {
        free(start);
        free(list);
        free(fill);
        return 0;
    }
    memcpy(fill, start, (size_t)nb * sizeof(int64_t));
    for (int64_t n = 0; n < npts; n++) {
        OMP_BANDS(n, b0, b1);
        for (int64_t b = b0; b <= b1; b++) list[fill[b]++] = n;
    }
#undef OMP_BANDS
    int used = 1;
#pragma omp parallel num_threads(nthreads)
    {
        int64_t nt = 1, tid = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        if (tid == 0) used = (int)nt;
        /* interleaved bands balance clustered inputs */
        for (int64_t b = tid; b < nb; b += nt) {
            const int64_t y0 = b * band, y1 = y0 + band < N ? y0 + band : N;
            grid_zz_list(out, uv, vals, list + start[b], start[b + 1] - start[b], N, gcf, S, r2c,
                         i2c, y0, y1);
        }
    }
    free(start);
    free(list);
    free(fill);
    return used;
}

/* gridCPU (grid.cu:87-139) on nthreads threads, bit-identical to
 * oracle_grid_zz with exact index math and a zeroed grid. */
int oracle_grid_zz_omp(double* out, const double* uv, const double* vals, int64_t npts,
                       int64_t N, const double* gcf, int64_t S, int nthreads)
{
    return grid_omp(out, uv, vals, npts, N, gcf, S, NULL, NULL, nthreads);
}

/* oracle_grid_const_tap (the COMPUTE_GCF form) on nthreads threads,
 * bit-identical to the serial loop. */
int oracle_grid_const_tap_omp(double* out, const double* uv, const double* vals, int64_t npts,
                              int64_t N, int64_t S, const float* r2, const float* i2,
                              int nthreads)
{
    return grid_omp(out, uv, vals, npts, N, NULL, S, r2, i2, nthreads);
}

/* PRECISION=float, OUTPRECISION=double: float inputs and float products,
 * summed into a double planar grid (grid.cu:169-185 with PRECISION float). */
void oracle_grid_zc(double* out, const float* uv, const float* vals, int64_t npts,
                    int64_t N, const float* gcf, int64_t S, int zero_first)
{
    if (zero_first) memset(out, 0, sizeof(double) * 2 * N * N);
    double* re = out;
    double* im = out + N * N;
    const int64_t h = S / 2;
    for (int64_t n = 0; n < npts; n++) {
        int mx, my, sx, sy;
        index_floorf((double)uv[2 * n], &mx, &sx);
        index_floorf((double)uv[2 * n + 1], &my, &sy);
        int64_t k = (int64_t)OVERSAMPLE * sy + sx;
        if (k < 0 || k >= OVERSAMPLE * OVERSAMPLE) continue;
        const float* g = gcf + 2 * k * S * S;
        const float r1 = vals[2 * n], i1 = vals[2 * n + 1];
        for (int64_t a = h; a > -h; a--) {
            const int64_t x = mx + a;
            if (x < 0 || x >= N) continue;
            for (int64_t b = h; b > -h; b--) {
                const int64_t y = my + b;
                if (y < 0 || y >= N) continue;
                const int64_t t = (b + h - 1) * S + (a + h - 1);
                const float r2 = g[2 * t], i2 = g[2 * t + 1];
                float pr = r1 * r2 - i1 * i2;
                float pi = r1 * i2 + r2 * i1;
                re[x + N * y] += (double)pr;
                im[x + N * y] += (double)pi;
            }
        }
    }
}

/* PRECISION=float, OUTPRECISION=float: the same with a float planar grid.
 * (The reference's gridCPU_pz casts its output to double* whatever its type,
 * grid.cu:153-154, so a float2 grid is written as doubles; the restatement
 * keeps float storage, which is what the GPU pair <float2,float2> produces.) */
void oracle_grid_cc(float* out, const float* uv, const float* vals, int64_t npts,
                    int64_t N, const float* gcf, int64_t S, int zero_first)
{
    if (zero_first) memset(out, 0, sizeof(float) * 2 * N * N);
    float* re = out;
    float* im = out + N * N;
    const int64_t h = S / 2;
    for (int64_t n = 0; n < npts; n++) {
        int mx, my, sx, sy;
        index_floorf((double)uv[2 * n], &mx, &sx);
        index_floorf((double)uv[2 * n + 1], &my, &sy);
        int64_t k = (int64_t)OVERSAMPLE * sy + sx;
        if (k < 0 || k >= OVERSAMPLE * OVERSAMPLE) continue;
        const float* g = gcf + 2 * k * S * S;
        const float r1 = vals[2 * n], i1 = vals[2 * n + 1];
        for (int64_t a = h; a > -h; a--) {
            const int64_t x = mx + a;
            if (x < 0 || x >= N) continue;
            for (int64_t b = h; b > -h; b--) {
                const int64_t y = my + b;
                if (y < 0 || y >= N) continue;
                const int64_t t = (b + h - 1) * S + (a + h - 1);
                const float r2 = g[2 * t], i2 = g[2 * t + 1];
                re[x + N * y] += r1 * r2 - i1 * i2;
                im[x + N * y] += r1 * i2 + r2 * i1;
            }
        }
    }
}

/* Number of coordinates whose gridCPU-form sub-pixel index (floorf on
 * float-rounded values, grid.cu:107-111) differs from the exact double one
 * the GPU kernels use (grid_gpu.cu:118-121); zero for float-representable uv. */
int64_t oracle_floorf_divergent(const double* uv, int64_t npts)
{
    int64_t cnt = 0;
    for (int64_t n = 0; n < 2 * npts; n++) {
        int m0, s0, m1, s1;
        index_exact(uv[n], &m0, &s0);
        index_floorf(uv[n], &m1, &s1);
        if (m0 != m1 || s0 != s1) cnt++;
    }
    return cnt;
}

/* ---- init_gcf (grid.cu:72-85), double precision ------------------------- */
void oracle_init_gcf(double* gcf, int64_t size)
{
    for (int64_t sub_x = 0; sub_x < OVERSAMPLE; sub_x++)
        for (int64_t sub_y = 0; sub_y < OVERSAMPLE; sub_y++)
            for (int64_t x = 0; x < size; x++)
                for (int64_t y = 0; y < size; y++) {
                    double tmp = sin(6.28 * x / size / OVERSAMPLE) *
                                 exp(-(1.0 * x * x + 1.0 * y * y * sub_y) / size / size / 2);
                    int64_t i = size * size * (sub_x + sub_y * OVERSAMPLE) + x + y * size;
                    gcf[2 * i] = tmp * sin(1.0 * x * sub_x / (y + 1)) + 0.4;
                    gcf[2 * i + 1] = tmp * cos(1.0 * x * sub_x / (y + 1)) - 0.2;
                }
}

/* ---- benchmark GCF: separable Gaussian x quadratic phase -----------------
 * G[k=8sy+sx][r][c]: dx = (c-S/2+1) - sx/8, dy = (r-S/2+1) - sy/8,
 * g = exp(-(dx^2+dy^2)/(2 sigma^2)) * exp(i*beta*(dx^2+dy^2)).            */
void oracle_config_gcf(double* gcf, int64_t S, double sigma, double beta)
{
    const int64_t h = S / 2;
    for (int64_t sy = 0; sy < OVERSAMPLE; sy++)
        for (int64_t sx = 0; sx < OVERSAMPLE; sx++)
            for (int64_t r = 0; r < S; r++)
                for (int64_t c = 0; c < S; c++) {
                    double dx = (double)(c - h + 1) - sx / 8.0;
                    double dy = (double)(r - h + 1) - sy / 8.0;
                    double q = dx * dx + dy * dy;
                    double amp = exp(-q / (2.0 * sigma * sigma));
                    int64_t i = ((sy * OVERSAMPLE + sx) * S + r) * S + c;
                    gcf[2 * i] = amp * cos(beta * q);
                    gcf[2 * i + 1] = amp * sin(beta * q);
                }
}

/* ---- generators ---------------------------------------------------------- */
/* SplitMix64 (Steele, Lea, Flood 2014; constants as published by Vigna). */
static inline uint64_t splitmix64(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double unif(uint64_t* s) /* [0,1) */
{
    return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}
static inline void box_muller(uint64_t* s, double* z0, double* z1)
{
    double u1 = 1.0 - unif(s); /* (0,1] */
    double u2 = unif(s);
    double r = sqrt(-2.0 * log(u1));
    *z0 = r * cos(6.283185307179586 * u2);
    *z1 = r * sin(6.283185307179586 * u2);
}

/* Uniform uv in [lo, hi) rounded to float32-representable doubles, values
 * complex N(0,1).  (Small-grid uniform workload.) */
void oracle_gen_uniform(double* uv, double* vals, int64_t npts, double lo, double hi,
                        uint64_t seed)
{
    uint64_t s = seed;
    for (int64_t n = 0; n < npts; n++) {
        uv[2 * n] = (double)(float)(lo + (hi - lo) * unif(&s));
        uv[2 * n + 1] = (double)(float)(lo + (hi - lo) * unif(&s));
        double z0, z1;
        box_muller(&s, &z0, &z1);
        vals[2 * n] = z0;
        vals[2 * n + 1] = z1;
    }
}

/* Radially clustered uv: angle U[0,2pi), radius Exp(mean), centred at
 * (cx, cy); any point whose S x S window would leave [0,N) is redrawn, so
 * every tap lands on the grid.  Values complex N(0,1). */
void oracle_gen_clustered(double* uv, double* vals, int64_t npts, int64_t N, int64_t S,
                          double mean_r, double cx, double cy, uint64_t seed)
{
    uint64_t s = seed;
    const int64_t h = S / 2;
    for (int64_t n = 0; n < npts; n++) {
        double u, v;
        for (;;) {
            double th = 6.283185307179586 * unif(&s);
            double r = -mean_r * log(1.0 - unif(&s));
            u = (double)(float)(cx + r * cos(th));
            v = (double)(float)(cy + r * sin(th));
            double fu = floor(u), fv = floor(v);
            if (fu - h + 1 >= 0 && fu + h <= N - 1 && fv - h + 1 >= 0 && fv + h <= N - 1) break;
        }
        uv[2 * n] = u;
        uv[2 * n + 1] = v;
        double z0, z1;
        box_muller(&s, &z0, &z1);
        vals[2 * n] = z0;
        vals[2 * n + 1] = z1;
    }
}

/* The reference's default input stream (grid.cu:444-452): glibc srand/rand,
 * all arithmetic in float, stored as double.  Called through the C library,
 * so the stream is the one the reference binary would draw. */
void oracle_ref_inputs(double* uv, double* vals, int64_t npts, int64_t img_size,
                       unsigned seed)
{
    srand(seed);
    for (int64_t n = 0; n < npts; n++) {
        uv[2 * n] = ((float)rand()) / RAND_MAX * img_size;
        uv[2 * n + 1] = ((float)rand()) / RAND_MAX * img_size;
        vals[2 * n] = ((float)rand()) / RAND_MAX;
        vals[2 * n + 1] = ((float)rand()) / RAND_MAX;
    }
}

/* ---- binning ------------------------------------------------------------- */
/* vis2ints (grid_gpu.cu:352-365): fixed point main*8+sub. */
void oracle_vis2ints(const double* uv, int64_t npts, int32_t* out)
{
    for (int64_t n = 0; n < npts; n++) {
        int mx, sx, my, sy;
        index_exact(uv[2 * n], &mx, &sx);
        index_exact(uv[2 * n + 1], &my, &sy);
        out[2 * n] = mx * OVERSAMPLE + sx;
        out[2 * n + 1] = my * OVERSAMPLE + sy;
    }
}

/* set_bookmarks (grid_gpu.cu:366-390) for cells of `blocksize` pixels,
 * blockgrid cells per row, on fixed-point input sorted by cell.
 * bookmarks[c] = first index whose cell >= c; bookmarks[blockgrid^2] = npts.
 * Restated without the reference's out-of-bounds read of in[npts] and its
 * overfill past the last cell. */
void oracle_set_bookmarks(const int32_t* ints, int64_t npts, int64_t blocksize,
                          int64_t blockgrid, int32_t* bookmarks)
{
    int64_t ncell = blockgrid * blockgrid;
    int64_t c = 0;
    for (int64_t q = 0; q < npts; q++) {
        int64_t cx = ints[2 * q] / OVERSAMPLE / blocksize;
        int64_t cy = ints[2 * q + 1] / OVERSAMPLE / blocksize;
        int64_t cell = cy * blockgrid + cx;
        while (c <= cell && c < ncell) bookmarks[c++] = (int32_t)q;
    }
    while (c <= ncell) bookmarks[c++] = (int32_t)npts;
}

/* The build's bin key.  q = main + S/2 is the last tap column plus one;
 * a visibility can touch the grid only if 0 <= q <= N+S-2 in both axes.
 * bin = q / H; key = (k*nby + by)*nbx + bx with nbx = nby = (N+S-2)/H + 1;
 * key = -1 for a visibility that touches no grid pixel. */
int64_t oracle_bin_key(double u, double v, int64_t N, int64_t S, int64_t H)
{
    int mx, sx, my, sy;
    index_exact(u, &mx, &sx);
    index_exact(v, &my, &sy);
    int64_t qx = (int64_t)mx + S / 2, qy = (int64_t)my + S / 2;
    int64_t qmax = N + S - 2;
    if (qx < 0 || qy < 0 || qx > qmax || qy > qmax) return -1;
    int64_t nb = qmax / H + 1;
    int64_t k = OVERSAMPLE * sy + sx;
    return (k * nb + qy / H) * nb + qx / H;
}

/* Stable counting sort of visibility indices by bin key.
 * perm[j] = index of the j-th visibility in key order (ties by index);
 * offsets[key] = CSR start, offsets[nkeys] = number of binned visibilities.
 * Returns that number (visibilities with key -1 are dropped). */
int64_t oracle_bin_sort(const double* uv, int64_t npts, int64_t N, int64_t S, int64_t H,
                        int32_t* perm, int32_t* offsets)
{
    int64_t nb = (N + S - 2) / H + 1;
    int64_t nkeys = (int64_t)OVERSAMPLE * OVERSAMPLE * nb * nb;
    int64_t* key = (int64_t*)malloc(sizeof(int64_t) * (npts > 0 ? npts : 1));
    memset(offsets, 0, sizeof(int32_t) * (nkeys + 1));
    for (int64_t n = 0; n < npts; n++) {
        key[n] = oracle_bin_key(uv[2 * n], uv[2 * n + 1], N, S, H);
        if (key[n] >= 0) offsets[key[n] + 1]++;
    }
    for (int64_t i = 0; i < nkeys; i++) offsets[i + 1] += offsets[i];
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * nkeys);
    memcpy(cur, offsets, sizeof(int32_t) * nkeys);
    for (int64_t n = 0; n < npts; n++)
        if (key[n] >= 0) perm[cur[key[n]]++] = (int32_t)n;
    int64_t total = offsets[nkeys];
    free(cur);
    free(key);
    return total;
}

/* ---- moving-window owner pixel (grid_gpu.cu:568-573) ---------------------
 * The unique pixel congruent to t (mod S) in [m-S/2+1, m+S/2], or -1 when the
 * reference skips it (m+S/2 < t, or the pixel is past the grid edge). */
int64_t oracle_window_pixel(int64_t m, int64_t t, int64_t S, int64_t N)
{
    int64_t h = S / 2;
    if (m + h < t) return -1;
    int64_t p = S * ((m + h - t) / S) + t;
    if (p >= N) return -1;
    return p;
}

/* ---- COMPUTE_GCF (grid_gpu.cu:103-108,137-147) ---------------------------
 * T = gcf_host[S(S-1)/2-1].x, w = .y; p1 = (float)(2*3.1415926*w); p2 = p1*T;
 * phase = p1 - p2*sqrt(dx^2+dy^2) in float with the scatter kernel's
 * dx = (float)(main_x - u + sub_x/8); the tap value is (sin, cos)(phase),
 * the same for every tap of the visibility. */
/* form 0: the scatter kernel's dx (float uv); form 1: the gather and
 * moving-window kernels' dx, computed from the int2 fixed point
 * inn = main*8+sub in place of u (grid_gpu.cu:456-459, :600-603). */
void oracle_compute_gcf_phase(const double* uv, int64_t npts, double T, double w, int form,
                              float* r2_out, float* i2_out)
{
    float p1 = (float)(2 * 3.1415926 * w);
    float p2 = (float)(p1 * T);
    for (int64_t n = 0; n < npts; n++) {
        int mx, sx, my, sy;
        index_exact(uv[2 * n], &mx, &sx);
        index_exact(uv[2 * n + 1], &my, &sy);
        double ux = uv[2 * n], uy = uv[2 * n + 1];
        if (form == 1) {
            ux = (double)(mx * OVERSAMPLE + sx);
            uy = (double)(my * OVERSAMPLE + sy);
        }
        float xs = (float)(mx - ux + sx * 1.0 / 8.0);
        float ys = (float)(my - uy + sy * 1.0 / 8.0);
        xs *= xs;
        ys *= ys;
        float ph = p1 - p2 * sqrtf(xs + ys);
        r2_out[n] = sinf(ph);
        i2_out[n] = cosf(ph);
    }
}

/* Grid with a per-visibility constant tap value (the COMPUTE_GCF form):
 * every in-grid tap of visibility n adds vals[n] * (r2[n] + i i2[n]); the
 * tap value is float (sincosf), the product and the sums are double, as in
 * the reference scatter kernel with CmplxType double2. */
void oracle_grid_const_tap(double* out, const double* uv, const double* vals, int64_t npts,
                           int64_t N, int64_t S, const float* r2, const float* i2)
{
    memset(out, 0, sizeof(double) * 2 * N * N);
    double* re = out;
    double* im = out + N * N;
    const int64_t h = S / 2;
    for (int64_t n = 0; n < npts; n++) {
        int mx, sx, my, sy;
        index_exact(uv[2 * n], &mx, &sx);
        index_exact(uv[2 * n + 1], &my, &sy);
        const double r1 = vals[2 * n], i1 = vals[2 * n + 1];
        const double pr = r1 * r2[n] - i1 * i2[n];
        const double pi = r1 * i2[n] + r2[n] * i1;
        for (int64_t a = h; a > -h; a--) {
            const int64_t x = mx + a;
            if (x < 0 || x >= N) continue;
            for (int64_t b = h; b > -h; b--) {
                const int64_t y = my + b;
                if (y < 0 || y >= N) continue;
                re[x + N * y] += pr;
                im[x + N * y] += pi;
            }
        }
    }
}

/* (u, v, w) in metres per row and channel frequencies in Hz -> grid pixel
 * coordinates, visibility n = row * nchan + chan (the order of the
 * reference's HDF5 reader: time outer, frequency inner, hdf52struct.cpp:95-113):
 *   u_pix = N/2 + (u f / c) N cell, v_pix likewise, c = 299792458 m/s;
 * with nplanes > 0 a w-plane GCF index
 *   g = min(nplanes - 1, floor(|w f / c| / w_max * nplanes)).
 * The reference's own normalised form, convert() (grid_gpu.cu:45-51), maps
 * pin in [-1, 1) to (pin + 1) * asize and splits it into main and sub the way
 * the gridder's index math does for non-negative coordinates. */
void oracle_uvw_to_pixels(const double* uvw, int64_t nrow, const double* freq, int64_t nchan,
                          int64_t N, double cell, int64_t nplanes, double w_max, double* uv,
                          int32_t* gidx)
{
    const double c = 299792458.0;
    for (int64_t r = 0; r < nrow; r++)
        for (int64_t ch = 0; ch < nchan; ch++) {
            const int64_t n = r * nchan + ch;
            const double s = freq[ch] / c;
            uv[2 * n] = 0.5 * (double)N + uvw[3 * r] * s * (double)N * cell;
            uv[2 * n + 1] = 0.5 * (double)N + uvw[3 * r + 1] * s * (double)N * cell;
            if (nplanes > 0 && gidx) {
                const double wl = fabs(uvw[3 * r + 2] * s);
                int64_t g = (int64_t)floor(wl / w_max * (double)nplanes);
                gidx[n] = (int32_t)(g < nplanes - 1 ? g : nplanes - 1);
            }
        }
}

/* Degridding, the adjoint of gridCPU (no counterpart in the reference):
 * vals[n] = sum over the in-grid taps of visibility n of
 *           conj(G[k][r][c]) * grid[main_y - S/2 + 1 + r][main_x - S/2 + 1 + c],
 * with the index math of gridCPU (exact floor).  grid is planar (2, N, N).
 * Pinned by <grid(v), g> = <v, degrid(g)> against oracle_grid_zz and by
 * single-pixel grids. */
void oracle_degrid_zz(double* vals, const double* uv, int64_t npts, const double* grid, int64_t N,
                      const double* gcf, int64_t S)
{
    const double* re = grid;
    const double* im = grid + N * N;
    const int64_t h = S / 2;
    for (int64_t n = 0; n < npts; n++) {
        int mx, sx, my, sy;
        index_exact(uv[2 * n], &mx, &sx);
        index_exact(uv[2 * n + 1], &my, &sy);
        const double* g = gcf + 2 * (int64_t)(8 * sy + sx) * S * S;
        double ar = 0.0, ai = 0.0;
        for (int64_t r = 0; r < S; r++) {
            const int64_t y = my - h + 1 + r;
            if (y < 0 || y >= N) continue;
            for (int64_t c = 0; c < S; c++) {
                const int64_t x = mx - h + 1 + c;
                if (x < 0 || x >= N) continue;
                const double gr = g[2 * (r * S + c)], gi = g[2 * (r * S + c) + 1];
                const double xr = re[x + N * y], xi = im[x + N * y];
                ar += gr * xr + gi * xi;
                ai += gr * xi - gi * xr;
            }
        }
        vals[2 * n] = ar;
        vals[2 * n + 1] = ai;
    }
}
