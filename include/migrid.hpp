/*
 * migrid.hpp -- header-only C++ drop-in for the reference's gridding calls.
 *
 *   template <class CmplxOutType, class CmplxType>
 *   void gridGPU(CmplxOutType* out, CmplxType* in, CmplxType* in_vals,
 *                size_t npts, size_t img_dim, CmplxType* gcf, size_t gcf_dim);
 *
 * has the parameter list of grid_gpu.cuh:30-32 and forwards to the C ABI in
 * migrid.h.  Any {x, y} complex type of 8 or 16 bytes (float2/double2 from
 * hip_runtime.h, or a plain struct) works; the three reference
 * instantiations <double2,double2>, <double2,float2>, <float2,float2>
 * (grid_gpu.cu:820-825) map to migrid_gridGPU_zz/_zc/_cc.
 *
 * The reference chooses its strategy with build flags (Makefile:36-47); this
 * header reads the same macros: __GATHER -> gather, __MOVING_WINDOW -> moving
 * window (planar output), otherwise scatter; __COMPUTE_GCF -> the reference's
 * on-the-fly GCF.  One flag of its own, MIGRID_USE_TILED, picks the tiled
 * strategy (the fastest, with no counterpart in the reference) where the
 * reference would scatter.  Errors are printed and execution continues, as the
 * reference's CUDA_CHECK_ERR does (grid_gpu.cu:32-36); use the C ABI to get
 * status codes instead.
 *
 * Defining MIGRID_DROPIN_GRIDCPU before including this header also defines
 * gridCPU / gridCPU_pz (grid.cu:87-193) with the same argument lists
 * (unpadded output base, planar layout), computed on the GPU.  With a float2
 * output, gridCPU_pz writes float planes (N^2 floats each): the reference
 * casts its output to double* whatever its type (grid.cu:142-144), writing
 * double planes past the end of an N^2 float2 buffer.
 */
#ifndef MIGRID_HPP
#define MIGRID_HPP

#include <cstddef>
#include <cstdio>

#include "migrid.h"

namespace migrid_dropin {

inline migrid_opts build_flag_opts()
{
    migrid_opts o;
    migrid_opts_default(&o);
#if defined(__GATHER)
    o.strategy = MIGRID_GATHER;
#elif defined(__MOVING_WINDOW)
    o.strategy = MIGRID_WINDOW;
#elif defined(MIGRID_USE_TILED)
    o.strategy = MIGRID_TILED;
#else
    o.strategy = MIGRID_SCATTER;
#endif
#if defined(__COMPUTE_GCF)
    o.gcf_mode = MIGRID_GCF_COMPAT;
#endif
    return o;
}

inline void report(int rc, const char* what)
{
    if (rc != MIGRID_OK) std::printf("Error %d in %s: %s\n", rc, what, migrid_last_error());
}

}  // namespace migrid_dropin

template <class CmplxOutType, class CmplxType>
void gridGPU(CmplxOutType* out, CmplxType* in, CmplxType* in_vals, size_t npts, size_t img_dim,
             CmplxType* gcf, size_t gcf_dim)
{
    static_assert(sizeof(CmplxType) == 8 || sizeof(CmplxType) == 16,
                  "CmplxType must be a {float,float} or {double,double} pair");
    static_assert(sizeof(CmplxOutType) == 8 || sizeof(CmplxOutType) == 16,
                  "CmplxOutType must be a {float,float} or {double,double} pair");
    static_assert(!(sizeof(CmplxOutType) == 8 && sizeof(CmplxType) == 16),
                  "the reference has no <float2,double2> instantiation");
    const migrid_opts o = migrid_dropin::build_flag_opts();
    int rc;
    if (sizeof(CmplxType) == 16)
        rc = migrid_gridGPU_zz((migrid_c128*)out, (const migrid_c128*)in,
                               (const migrid_c128*)in_vals, npts, img_dim,
                               (const migrid_c128*)gcf, gcf_dim, &o);
    else if (sizeof(CmplxOutType) == 16)
        rc = migrid_gridGPU_zc((migrid_c128*)out, (const migrid_c64*)in,
                               (const migrid_c64*)in_vals, npts, img_dim, (const migrid_c64*)gcf,
                               gcf_dim, &o);
    else
        rc = migrid_gridGPU_cc((migrid_c64*)out, (const migrid_c64*)in,
                               (const migrid_c64*)in_vals, npts, img_dim, (const migrid_c64*)gcf,
                               gcf_dim, &o);
    migrid_dropin::report(rc, "gridGPU");
}

#ifdef MIGRID_DROPIN_GRIDCPU
/* gridCPU: zeroes the N*N complex output, planar result (grid.cu:87-139). */
template <class CmplxType>
void gridCPU(CmplxType* out, CmplxType* in, CmplxType* in_vals, size_t npts, size_t img_dim,
             CmplxType* gcf, size_t gcf_dim)
{
    static_assert(sizeof(CmplxType) == 16, "gridCPU is provided for double2 data");
    migrid_dropin::report(
        migrid_gridCPU_zz((migrid_c128*)out, (const migrid_c128*)in, (const migrid_c128*)in_vals,
                          npts, img_dim, (const migrid_c128*)gcf, gcf_dim),
        "gridCPU");
}

/* gridCPU_pz: adds into the planar output (grid.cu:140-193). */
template <class CmplxOutType, class CmplxType>
void gridCPU_pz(CmplxOutType* out, CmplxType* in, CmplxType* in_vals, size_t npts,
                size_t img_dim, CmplxType* gcf, size_t gcf_dim)
{
    int rc;
    if (sizeof(CmplxType) == 16)
        rc = migrid_gridCPU_pz_zz((migrid_c128*)out, (const migrid_c128*)in,
                                  (const migrid_c128*)in_vals, npts, img_dim,
                                  (const migrid_c128*)gcf, gcf_dim);
    else if (sizeof(CmplxOutType) == 16)
        rc = migrid_gridCPU_pz_zc((migrid_c128*)out, (const migrid_c64*)in,
                                  (const migrid_c64*)in_vals, npts, img_dim,
                                  (const migrid_c64*)gcf, gcf_dim);
    else
        rc = migrid_gridCPU_pz_cc((migrid_c64*)out, (const migrid_c64*)in,
                                  (const migrid_c64*)in_vals, npts, img_dim,
                                  (const migrid_c64*)gcf, gcf_dim);
    migrid_dropin::report(rc, "gridCPU_pz");
}
#endif

#endif /* MIGRID_HPP */
