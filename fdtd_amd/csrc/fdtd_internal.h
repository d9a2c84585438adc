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
    // Scaled-unit form for 16-bit storage (fields stored as E/sE, (D/eps0)/sD,
    // (eta0 H)/sH; all arithmetic fp32):
    // E' = k1 E - k2 E_prev + k3 (D' - D) + k5 ((D' - D) - (D - D_prev)).
    float k1, k2, k3, k5;
};

// Material id byte: low 7 bits index the coefficient table, the top bit marks
// a soft Jz source cell (fdtd_calculations.cu:187-213 semantics).
constexpr int kMaxMaterials = 128;
constexpr uint8_t kSourceBit = 0x80;
constexpr uint8_t kMaterialMask = 0x7f;

// Host-side coefficient builder shared by the legacy and fused paths.
MatCoef fdtd_mat_coef(float dt, float eps0, float sigma, float epsS, float epsI, float tau, int debye_mode);

// Field storage of a context.
enum Storage { kStoreF32 = 0, kStoreF16 = 1, kStoreBF16 = 2, kStoreF32Fast = 3 };

// Arguments of the fused H -> D (+source) -> E kernel for one time step.
// Field pointers point at element type float, __half or __hip_bfloat16
// according to the launch's storage.
struct StepArgs {
    const void *hxo, *hyo, *hzo;    // H^{n-1/2}
    void *hxn, *hyn, *hzn;          // H^{n+1/2}
    const void *ex0, *ey0, *ez0;    // E^n
    const void *ex1, *ey1, *ez1;    // E^{n-1}
    void *exT, *eyT, *ezT;          // E^{n+1} (ring target slot)
    const void *dx0, *dy0, *dz0;    // D^n
    const void *dx1, *dy1, *dz1;    // D^{n-1}
    void *dxT, *dyT, *dzT;          // D^{n+1} (ring target slot)
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
    void *exT, *eyT, *ezT;               // E^{n+1}
    const void *exS, *eyS, *ezS;         // E^n
    const float *rpx0, *rpy0, *rpz0, *rpxEnd, *rpyEnd, *rpzEnd;
    int nx, ny, nz, pitch;
    long long plane;
    float dt, hx, hy, hz, mu0, eps0;
    int seg_start[13];                   // prefix sums of the 12 face segments
};

struct ProbeArgs {
    const void *field[9];                // Dx Dy Dz Hx Hy Hz Ex Ey Ez of the newest level
    double factor[9];                    // stored value -> SI (1 for fp32 storage)
    const int *probes;                   // nprobe x (field, i, j, k)
    int nprobe;
    int pitch;
    long long plane;
    double *out;                         // [step][probe]
    int row;
};

hipError_t launch_fused_step(const StepArgs &a, int ty, int storage, hipStream_t s);
hipError_t launch_mur(const MurArgs &a, int storage, hipStream_t s);
hipError_t launch_probes(const ProbeArgs &a, int storage, hipStream_t s);
void mur_segments(int nx, int ny, int nz, int *seg_start);
