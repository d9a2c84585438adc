/*
 * fdtd.h -- state and parameter types of the FDTD path and the host-side
 * services that set it up, exported with C linkage by libfdtd_amd.so.
 *
 * Drop-in for the reference's fdtd.h: FdtdParams and FdtdField keep the
 * reference's field order and types byte for byte (fdtd.h:7-34, fdtd.h:37-54),
 * so code that fills or reads them is unchanged.  CopyParams/ResultsParams
 * (fdtd.h:57-82) carry a hipStream_t where the reference has cudaStream_t.
 *
 * Error behaviour follows the reference's CHECK macro (utils.h:5-13): the
 * host services print "Error: <file>: <line>" and "Code <n>: <message>" and
 * exit(EXIT_FAILURE) when a HIP call fails, and print and exit when an input
 * file cannot be opened (fdtd.cu:889-892, 914-917).  initParamsWithPath also
 * exits on an unreadable params file, where the reference crashes
 * (fdtd.cu:553-554).
 */
#ifndef FDTD_H
#define FDTD_H

#include <hip/hip_runtime_api.h>
#include <pthread.h>

#ifdef __cplusplus
extern "C" {
#endif

/* fdtd.h:7-34 */
typedef struct
{
    int   nx, ny, nz;
    int   iterationsCount;
    char  *inputPath;
    char  *outputPath;
    int   elementsPerWave;
    float waveFrequency;
    float pulseWidth;
    float pulseModulationFrequency;
    int   sourcesCount;
    int   *sources;
    float defaultSigma;
    float defaultEpsS;
    float defaultEpsI;
    float defaultTauD;
    int   jzCount;
    float *jz;

    float pi;
    float c;
    float timeskip;
    float lambda;
    float dt;
    float dx, dy, dz;
    float mu0;
    float eps0;
} FdtdParams;

/* fdtd.h:37-54: three ring slots of E and D, one H, four material arrays
 * and six two-cell-thick Mur permittivity planes. */
typedef struct
{
    float *ex0, *ey0, *ez0;
    float *ex1, *ey1, *ez1;
    float *ex2, *ey2, *ez2;

    float *hx, *hy, *hz;

    float *dx0, *dy0, *dz0;
    float *dx1, *dy1, *dz1;
    float *dx2, *dy2, *dz2;

    float *epsI, *epsS;
    float *tauD, *sigma;

    float *rpx0,   *rpy0,   *rpz0;
    float *rpxEnd, *rpyEnd, *rpzEnd;
} FdtdField;

/* fdtd.h:57-71 */
typedef struct
{
    float *xSource;
    float *ySource;
    float *zSource;

    float *xBuffer;
    float *yBuffer;
    float *zBuffer;

    FdtdParams *params;
    hipStream_t stream;
    pthread_mutex_t *copyMutex;
} CopyParams;

/* fdtd.h:74-82 */
typedef struct
{
    FdtdParams *params;

    CopyParams *hParams;
    CopyParams *dParams;
    CopyParams *eParams;

    int currentIteration;
} ResultsParams;

/* Replaces initParamsWithPath, fdtd.cu:547-637: positional parse of the
 * input-parameter file and the derived constants.  Extension keys appended
 * after the last line the reference reads (ftau_d) are parsed too; see
 * fdtd_amd.h (FdtdExt). */
FdtdParams *initParamsWithPath(const char *filePath);
/* fdtd.cu:640-645 */
void deallocParams(FdtdParams *params);
/* fdtd.cu:648-674 */
void printParams(FdtdParams *params);

/* fdtd.cu:677-725: pinned host fields, material arrays set to the defaults,
 * host rp planes. */
FdtdField *initFieldWithParams(FdtdParams *params);
/* fdtd.cu:728-761 */
void deallocField(FdtdField *field);

/* fdtd.cu:764-822: 31 device arrays. */
FdtdField *initDeviceFieldWithParams(FdtdParams *params);
/* fdtd.cu:825-872 */
void deallocDeviceField(FdtdField *field);

/* fdtd.cu:942-1024 */
void setupMurBoundary(FdtdParams *params, FdtdField *field);
/* fdtd.cu:875-939 */
void loadMaterials(FdtdParams *params, FdtdField *field, const char *specsFilePath, const char *materialsPath);
/* fdtd.cu:1027-1082 */
void setupSources(FdtdParams *params);
/* fdtd.cu:1085-1133 */
void copyDataToDevice(FdtdParams *params, FdtdField *field, FdtdField *deviceField);
/* fdtd.cu:1136-1160 */
void *copyResultsWithParams(void *params);
/* fdtd.cu:1163-1194 */
void *writeResultsWithParams(void *params);
/* fdtd.cu:1197-1285: line probes through every source (1-based iteration in
 * the file name). */
void writeResults(FdtdParams *params,
                  float *hxSource, float *hySource, float *hzSource,
                  float *dxSource, float *dySource, float *dzSource,
                  float *exSource, float *eySource, float *ezSource,
                  int currentIteration);
/* fdtd.cu:1288-1494: every z slice of all nine fields as text. */
void writeResults3d(FdtdParams *params,
                    float *hxSource, float *hySource, float *hzSource,
                    float *dxSource, float *dySource, float *dzSource,
                    float *exSource, float *eySource, float *ezSource,
                    int currentIteration);

#ifdef __cplusplus
}
#endif

#endif
