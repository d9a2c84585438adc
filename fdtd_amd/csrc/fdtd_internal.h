// Internal declarations shared by the HIP translation units of libfdtd_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fdtd_amd.h"

// Per-material Debye coefficients in the precision the reference evaluates
// them (fdtd_calculations.cu:141-155): a, c, d are double expressions rounded
// to float; the three D weights are double.  With these, an E update
// c*(d*E^n - a*E^{n-1} + g1*D^{n+1} - g2*D^n + g3*D^{n-1}) evaluated as
// float products, then double sums, reproduces the reference's rounding.
struct MatCoef {
    float a, c, d, pad;
    double g1, g2, g3;
};

// Material id byte: low 7 bits index the coefficient table, the top bit marks
// a soft Jz source cell (fdtd_calculations.cu:187-213 semantics).
constexpr int kMaxMaterials = 128;
constexpr uint8_t kSourceBit = 0x80;
constexpr uint8_t kMaterialMask = 0x7f;

// Host-side coefficient builder shared by the legacy and fused paths.
MatCoef fdtd_mat_coef(float dt, float eps0, float sigma, float epsS, float epsI, float tau, int debye_mode);

// Arguments of the fused H -> D (+source) -> E kernel for one time step.
struct StepArgs {
    const float *hxo, *hyo, *hzo;   // H^{n-1/2}
    float *hxn, *hyn, *hzn;         // H^{n+1/2}
    const float *ex0, *ey0, *ez0;   // E^n
    const float *ex1, *ey1, *ez1;   // E^{n-1}
    float *exT, *eyT, *ezT;         // E^{n+1} (ring target slot)
    const float *dx0, *dy0, *dz0;   // D^n
    const float *dx1, *dy1, *dz1;   // D^{n-1}
    float *dxT, *dyT, *dzT;         // D^{n+1} (ring target slot)
    const uint8_t *ids;
    const MatCoef *lut;
    int nmat;
    float jz;
    int nx, ny, nz;
    int pitch;                      // elements between rows (>= nx)
    long long plane;                // elements between z planes (>= pitch*ny)
    float chx, chy, chz;            // dt/(mu0*dx), dt/(mu0*dy), dt/(mu0*dz)
    float cdx, cdy, cdz;            // dt/dx, dt/dy, dt/dz
    int kchunk;                     // z planes per block
    int tiles_x, tiles_y, tiles_z;  // logical tile grid
    int xcd_remap;                  // 0: identity, 1: contiguous runs per XCD, 2: layer-interleaved
    unsigned guard;                 // addressable elements before and after every field array (and ids)
};

// Guard (in elements) the fused kernel needs around every field array and
// the id array so that its halo loads need no bounds checks: one plane, ten
// rows and 256 elements (tiles are at most 8 rows high and 64 wide).
inline long long fused_guard(int pitch, long long plane) { return plane + 10LL * pitch + 256; }

// Arguments of the 6-face first-order Mur kernel (fdtd_calculations.cu:216-428).
struct MurArgs {
    float *exT, *eyT, *ezT;              // E^{n+1}
    const float *exS, *eyS, *ezS;        // E^n
    const float *rpx0, *rpy0, *rpz0, *rpxEnd, *rpyEnd, *rpzEnd;
    int nx, ny, nz, pitch;
    long long plane;
    float dt, hx, hy, hz, mu0, eps0;
    int seg_start[13];                   // prefix sums of the 12 face segments
};

struct ProbeArgs {
    const float *field[9];               // Dx Dy Dz Hx Hy Hz Ex Ey Ez of the newest level
    const int *probes;                   // nprobe x (field, i, j, k)
    int nprobe;
    int pitch;
    long long plane;
    double *out;                         // [step][probe]
    int row;
};

hipError_t launch_fused_step(const StepArgs &a, int ty, hipStream_t s);
hipError_t launch_mur(const MurArgs &a, hipStream_t s);
hipError_t launch_probes(const ProbeArgs &a, hipStream_t s);
void mur_segments(int nx, int ny, int nz, int *seg_start);
