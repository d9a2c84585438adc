// Drop-in device calculation entry points (fdtd_calculations.h).
//
// These replace the reference's five __global__ kernels and
// copySymbolsToDevice (fdtd_calculations.cu:7-428) one for one, with the same
// pointer lists, ranges and arithmetic (float products and sums, the double
// promotions of the Debye update, float Mur with correctly rounded sqrt and
// division).  Built with -ffp-contract=off, they are bit-identical to a
// straight C evaluation of the reference's equations.
//
// Launch geometry is the library's: one lane per cell over the flattened
// grid, 256 lanes per block, so every lane of a wave64 is busy whatever nx is
// (the reference's (128,1,1) blocks leave half the lanes idle at nx = 64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "fdtd_calculations.h"
#include "fdtd_internal.h"

namespace {

struct LegacyParams {
    int nx, ny, nz;
    float dt, dx, dy, dz, mu0, eps0;
    int nsrc;
    int njz;
    int debye_mode;
};

__constant__ LegacyParams c_p;
__constant__ int c_sources[D_MAX_SOURCES * 3];
__constant__ float c_jz[D_MAX_JZ];

LegacyParams h_p = {0, 0, 0, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0, 0, FDTD_DEBYE_FIXED};
bool h_params_set = false;

constexpr int kBlock = 256;

__device__ inline bool cell_of(long long q, const LegacyParams &p, int &i, int &j, int &k)
{
    const long long nxy = (long long)p.nx * p.ny;
    if (q >= nxy * p.nz) return false;
    k = (int)(q / nxy);
    const long long r = q - (long long)k * nxy;
    j = (int)(r / p.nx);
    i = (int)(r - (long long)j * p.nx);
    return true;
}

// fdtd_calculations.cu:32-77
__global__ __launch_bounds__(kBlock) void legacy_h(float *hx, float *hy, float *hz, const float *ex, const float *ey,
                                                   const float *ez)
{
    const LegacyParams p = c_p;
    int i, j, k;
    const long long o = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (!cell_of(o, p, i, j, k)) return;
    const long long sy = p.nx, sz = (long long)p.nx * p.ny;
    if (i >= 1 && i < p.nx - 1 && j < p.ny - 1 && k < p.nz - 1)
        hx[o] = hx[o] - p.dt / (p.mu0 * p.dy) * (ez[o + sy] - ez[o]) + p.dt / (p.mu0 * p.dy) * (ey[o + sz] - ey[o]);
    if (i < p.nx - 1 && j >= 1 && j < p.ny - 1 && k < p.nz - 1)
        hy[o] = hy[o] - p.dt / (p.mu0 * p.dz) * (ex[o + sz] - ex[o]) + p.dt / (p.mu0 * p.dx) * (ez[o + 1] - ez[o]);
    if (i < p.nx - 1 && j < p.ny - 1 && k >= 1 && k < p.nz - 1)
        hz[o] = hz[o] - p.dt / (p.mu0 * p.dx) * (ey[o + 1] - ey[o]) + p.dt / (p.mu0 * p.dy) * (ex[o + sy] - ex[o]);
}

// fdtd_calculations.cu:80-122
__global__ __launch_bounds__(kBlock) void legacy_d(float *dxT, float *dyT, float *dzT, const float *dxS,
                                                   const float *dyS, const float *dzS, const float *hx,
                                                   const float *hy, const float *hz)
{
    const LegacyParams p = c_p;
    int i, j, k;
    const long long o = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (!cell_of(o, p, i, j, k)) return;
    const long long sy = p.nx, sz = (long long)p.nx * p.ny;
    const float x = p.dt / p.dx, y = p.dt / p.dy, z = p.dt / p.dz;
    if (i < p.nx - 1 && j >= 1 && j < p.ny - 1 && k >= 1 && k < p.nz - 1)
        dxT[o] = dxS[o] + y * (hz[o] - hz[o - sy]) - z * (hy[o] - hy[o - sz]);
    if (i >= 1 && i < p.nx - 1 && j < p.ny - 1 && k >= 1 && k < p.nz - 1)
        dyT[o] = dyS[o] + z * (hx[o] - hx[o - sz]) - x * (hz[o] - hz[o - 1]);
    if (i >= 1 && i < p.nx - 1 && j >= 1 && j < p.ny - 1 && k < p.nz - 1)
        dzT[o] = dzS[o] + x * (hy[o] - hy[o - 1]) - y * (hx[o] - hx[o - sy]);
}

// fdtd_calculations.cu:187-213: one lane per listed source.
__global__ void legacy_sources(float *dzT, const float *dzS, const float *hx, const float *hy, int it)
{
    const LegacyParams p = c_p;
    const int s = threadIdx.x;
    if (s >= p.nsrc) return;
    const int x = c_sources[3 * s], y = c_sources[3 * s + 1], z = c_sources[3 * s + 2];
    const long long o = ((long long)z * p.ny + y) * p.nx + x;
    const float jz = (it >= 0 && it < p.njz) ? c_jz[it] : 0.0f;
    const float val = dzS[o] + p.dt / p.dx * (hy[o] - hy[o - 1]) - p.dt / p.dy * (hx[o] - hx[o - p.nx]) - jz;
    dzT[o] = val;
}

__device__ inline float e_update(float c, float d, float a, double g1, double g2, double g3, float e0, float e1,
                                 float d0, float d1, float d2)
{
    // c * (d*E^n - a*E^{n-1} + g1*D^{n+1} - g2*D^n + g3*D^{n-1}): the first
    // two products and their difference in float, the rest in double.
    const float fe = d * e0 - a * e1;
    double s = (double)fe + g1 * (double)d0;
    s = s - g2 * (double)d1;
    s = s + g3 * (double)d2;
    return (float)((double)c * s);
}

// fdtd_calculations.cu:125-184
__global__ __launch_bounds__(kBlock) void legacy_e(float *exT, float *eyT, float *ezT, const float *ex0,
                                                   const float *ey0, const float *ez0, const float *ex1,
                                                   const float *ey1, const float *ez1, const float *dx0,
                                                   const float *dy0, const float *dz0, const float *dx1,
                                                   const float *dy1, const float *dz1, const float *dx2,
                                                   const float *dy2, const float *dz2, const float *sigma,
                                                   const float *epsI, const float *epsS, const float *tauD)
{
    const LegacyParams p = c_p;
    int i, j, k;
    const long long o = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (!cell_of(o, p, i, j, k)) return;
    const bool ux = (i < p.nx - 1 && j >= 1 && j < p.ny - 1 && k >= 1 && k < p.nz - 1);
    const bool uyz = (i >= 1 && k >= 1);
    if (!ux && !uyz) return;
    const float sg = sigma[o], ei = epsI[o], es = epsS[o], tau = tauD[o];
    const float dt = p.dt, e0 = p.eps0;
    const float a = 2.0 * e0 * ei * tau;
    const float b = a + 2.0 * dt * (e0 * es + sg * tau);
    const float c = 1.0 / (b + sg * dt * dt);
    const float d = (p.debye_mode == FDTD_DEBYE_REF) ? (float)(2.0 * b - sg * dt * dt)
                                                      : (float)((double)a + (double)b - (double)(sg * dt * dt));
    const double g1 = 2.0 * (dt + tau);
    const double g2 = 2.0 * dt + 4.0 * tau;
    const double g3 = 2.0 * tau;
    if (ux) exT[o] = e_update(c, d, a, g1, g2, g3, ex0[o], ex1[o], dx0[o], dx1[o], dx2[o]);
    if (uyz) {
        eyT[o] = e_update(c, d, a, g1, g2, g3, ey0[o], ey1[o], dy0[o], dy1[o], dy2[o]);
        ezT[o] = e_update(c, d, a, g1, g2, g3, ez0[o], ez1[o], dz0[o], dz1[o], dz2[o]);
    }
}

inline int set_err(hipError_t e) { return (int)e; }

}  // namespace

extern "C" int fdtdSetDebyeMode(int mode)
{
    if (mode != FDTD_DEBYE_REF && mode != FDTD_DEBYE_FIXED) return FDTD_EINVAL;
    h_p.debye_mode = mode;
    if (h_params_set) return set_err(hipMemcpyToSymbol(HIP_SYMBOL(c_p), &h_p, sizeof(h_p)));
    return 0;
}

extern "C" int copySymbolsToDevice(FdtdParams *params)
{
    if (!params || params->nx < 2 || params->ny < 2 || params->nz < 2) return FDTD_EINVAL;
    const long long n = (long long)params->nx * params->ny * params->nz;
    if (n >= (1LL << 40)) return FDTD_ERANGE;
    h_p.nx = params->nx; h_p.ny = params->ny; h_p.nz = params->nz;
    h_p.dt = params->dt; h_p.dx = params->dx; h_p.dy = params->dy; h_p.dz = params->dz;
    h_p.mu0 = params->mu0; h_p.eps0 = params->eps0;
    const int ns = std::max(0, std::min(D_MAX_SOURCES, params->sourcesCount));
    const int nj = std::max(0, std::min(D_MAX_JZ, params->jzCount));
    for (int s = 0; s < ns; s++)
        for (int q = 0; q < 3; q++) {
            const int v = params->sources[3 * s + q];
            const int lim = q == 0 ? params->nx : (q == 1 ? params->ny : params->nz);
            // the source kernel reads x-1 and y-1
            if (v < (q < 2 ? 1 : 0) || v >= lim) return FDTD_ERANGE;
        }
    h_p.nsrc = ns;
    h_p.njz = params->jz ? nj : 0;
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_p), &h_p, sizeof(h_p));
    if (e == hipSuccess && ns > 0) e = hipMemcpyToSymbol(HIP_SYMBOL(c_sources), params->sources, sizeof(int) * 3 * ns);
    if (e == hipSuccess && h_p.njz > 0) e = hipMemcpyToSymbol(HIP_SYMBOL(c_jz), params->jz, sizeof(float) * h_p.njz);
    if (e == hipSuccess) h_params_set = true;
    return set_err(e);
}

static inline dim3 flat_grid()
{
    const long long n = (long long)h_p.nx * h_p.ny * h_p.nz;
    return dim3((unsigned)((n + kBlock - 1) / kBlock));
}

extern "C" int updateHField(float *hx, float *hy, float *hz, float *exSource, float *eySource, float *ezSource,
                            hipStream_t stream)
{
    if (!h_params_set) return FDTD_ESTATE;
    legacy_h<<<flat_grid(), kBlock, 0, stream>>>(hx, hy, hz, exSource, eySource, ezSource);
    return set_err(hipGetLastError());
}

extern "C" int updateDField(float *dxTarget, float *dyTarget, float *dzTarget, float *dxSource, float *dySource,
                            float *dzSource, float *hx, float *hy, float *hz, hipStream_t stream)
{
    if (!h_params_set) return FDTD_ESTATE;
    legacy_d<<<flat_grid(), kBlock, 0, stream>>>(dxTarget, dyTarget, dzTarget, dxSource, dySource, dzSource, hx, hy,
                                                 hz);
    return set_err(hipGetLastError());
}

extern "C" int updateEField(float *exTarget, float *eyTarget, float *ezTarget, float *exSource0, float *eySource0,
                            float *ezSource0, float *exSource1, float *eySource1, float *ezSource1, float *dxSource0,
                            float *dySource0, float *dzSource0, float *dxSource1, float *dySource1, float *dzSource1,
                            float *dxSource2, float *dySource2, float *dzSource2, float *sigma, float *epsI,
                            float *epsS, float *tauD, hipStream_t stream)
{
    if (!h_params_set) return FDTD_ESTATE;
    legacy_e<<<flat_grid(), kBlock, 0, stream>>>(exTarget, eyTarget, ezTarget, exSource0, eySource0, ezSource0,
                                                 exSource1, eySource1, ezSource1, dxSource0, dySource0, dzSource0,
                                                 dxSource1, dySource1, dzSource1, dxSource2, dySource2, dzSource2,
                                                 sigma, epsI, epsS, tauD);
    return set_err(hipGetLastError());
}

extern "C" int updateSources(float *dzTarget, float *dzSource, float *hx, float *hy, int currIteration,
                             hipStream_t stream)
{
    if (!h_params_set) return FDTD_ESTATE;
    if (h_p.nsrc == 0) return 0;
    legacy_sources<<<1, 64, 0, stream>>>(dzTarget, dzSource, hx, hy, currIteration);
    return set_err(hipGetLastError());
}

extern "C" int updateMurBoundary(float *exTarget, float *eyTarget, float *ezTarget, float *exSource,
                                 float *eySource, float *ezSource, float *rpx0, float *rpy0, float *rpz0,
                                 float *rpxEnd, float *rpyEnd, float *rpzEnd, hipStream_t stream)
{
    if (!h_params_set) return FDTD_ESTATE;
    MurArgs m;
    m.exT = exTarget; m.eyT = eyTarget; m.ezT = ezTarget;
    m.exS = exSource; m.eyS = eySource; m.ezS = ezSource;
    m.rpx0 = rpx0; m.rpy0 = rpy0; m.rpz0 = rpz0; m.rpxEnd = rpxEnd; m.rpyEnd = rpyEnd; m.rpzEnd = rpzEnd;
    m.nx = h_p.nx; m.ny = h_p.ny; m.nz = h_p.nz; m.pitch = h_p.nx; m.plane = (long long)h_p.nx * h_p.ny;
    m.dt = h_p.dt; m.hx = h_p.dx; m.hy = h_p.dy; m.hz = h_p.dz; m.mu0 = h_p.mu0; m.eps0 = h_p.eps0;
    mur_segments(h_p.nx, h_p.ny, h_p.nz, m.seg_start);
    return set_err(launch_mur(m, kStoreF32, stream));
}
