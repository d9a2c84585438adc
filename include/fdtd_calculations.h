/*
 * fdtd_calculations.h -- the device calculation entry points of the FDTD
 * time step, exported with C linkage by libfdtd_amd.so.
 *
 * Drop-in for the reference's fdtd_calculations.h:18-42.  The names and
 * pointer lists are the reference's; the differences are:
 *   * each update is a host function with a trailing hipStream_t instead of a
 *     __global__ kernel launched with <<<grid, block, 0, stream>>>; the
 *     library picks the launch geometry;
 *   * each returns an int (a hipError_t value, 0 = success) instead of void;
 *     C callers that ignore it stay source-compatible;
 *   * the parameter block is a library-private __constant__ copy, so the
 *     header defines no device symbols.
 * The limits D_MAX_SOURCES and D_MAX_JZ are the reference's
 * (fdtd_calculations.h:7-8); jz entries past the uploaded table read as 0
 * (the reference reads past its 512-entry constant array).
 */
#ifndef FDTD_CALCULATIONS_H
#define FDTD_CALCULATIONS_H

#include <hip/hip_runtime_api.h>

#include "fdtd.h"

#define D_MAX_SOURCES 16
#define D_MAX_JZ 512

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces copySymbolsToDevice, fdtd_calculations.cu:7-29: uploads nx/ny/nz,
 * dt, dx/dy/dz, mu0, eps0, min(sourcesCount, 16) sources and
 * min(jzCount, 512) jz values. */
int copySymbolsToDevice(FdtdParams *params);

/* Replaces __global__ updateHField, fdtd_calculations.cu:32-77 (in place). */
int updateHField(float *hx,       float *hy,       float *hz,
                 float *exSource, float *eySource, float *ezSource,
                 hipStream_t stream);

/* Replaces __global__ updateDField, fdtd_calculations.cu:80-122. */
int updateDField(float *dxTarget, float *dyTarget, float *dzTarget,
                 float *dxSource, float *dySource, float *dzSource,
                 float *hx,       float *hy,       float *hz,
                 hipStream_t stream);

/* Replaces __global__ updateEField, fdtd_calculations.cu:125-184.  The Debye
 * d coefficient follows fdtdSetDebyeMode (default FDTD_DEBYE_FIXED,
 * d = a + b - sigma dt^2; FDTD_DEBYE_REF reproduces the reference's
 * d = 2b - sigma dt^2 of fdtd_calculations.cu:144). */
int updateEField(float *exTarget,  float *eyTarget,  float *ezTarget,
                 float *exSource0, float *eySource0, float *ezSource0,
                 float *exSource1, float *eySource1, float *ezSource1,
                 float *dxSource0, float *dySource0, float *dzSource0,
                 float *dxSource1, float *dySource1, float *dzSource1,
                 float *dxSource2, float *dySource2, float *dzSource2,
                 float *sigma, float *epsI, float *epsS, float *tauD,
                 hipStream_t stream);

/* Replaces __global__ updateSources, fdtd_calculations.cu:187-213. */
int updateSources(float *dzTarget, float *dzSource,
                  float *hx,       float *hy,
                  int currIteration,
                  hipStream_t stream);

/* Replaces __global__ updateMurBoundary, fdtd_calculations.cu:216-428. */
int updateMurBoundary(float *exTarget, float *eyTarget, float *ezTarget,
                      float *exSource, float *eySource, float *ezSource,
                      float *rpx0,     float *rpy0,     float *rpz0,
                      float *rpxEnd,   float *rpyEnd,   float *rpzEnd,
                      hipStream_t stream);

/* Debye mode of updateEField: FDTD_DEBYE_FIXED (1, default) or
 * FDTD_DEBYE_REF (0).  No counterpart in the reference, which has only the
 * unstable form. */
int fdtdSetDebyeMode(int mode);

#ifdef __cplusplus
}
#endif

#endif
