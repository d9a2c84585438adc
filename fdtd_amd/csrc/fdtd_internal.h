// Internal declarations shared by the HIP translation units of libfdtd_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <algorithm>
#include <initializer_list>
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
    int xcd_remap;                  // block -> tile order (FdtdExt.xcd_remap; xcd_logical)
    unsigned guard;                 // addressable elements before and after every field array (and ids)
    // Mur x faces updated inside the fused kernel (fused_mur_x): 0 none, 1
    // all four (Ey, Ez at i = 0 and nx-1), 2 the two at i = 0 only.
    int mur_x;
    const float4 *murx0, *murx1;    // per (k, j) coefficients of the x-low / x-high faces (mur_x_coefs)
    int cpl;                        // cells per lane: 1, or 2 for 16-bit storage (128-cell tiles)
    int txw;                        // threads per tile row with one cell per lane: 64 or 128
    // Exact fp32 form: a half wave (32 cells of a tile row) whose cells all
    // have materials whose Debye update multiplies E^{n-1} and D^{n-1} by +0
    // (a = +0 and g3 = +0, i.e. tau = +0) does not load those six operands
    // (see fused_tile).
    int zskip;
};

// Coefficients {q, A, B, C} of the first-order Mur update of one x-face
// cell, E' = q*((A*tn + B*sn) - C*sf), evaluated in float exactly as the
// reference's expression (fdtd_calculations.cu:303-329, 369-395) rounds
// them: q = 1/(dt + h sqrt(mu0 eps0 rp_face)), A = dt - h sqrt(mu0 eps0
// rp_nb), B = dt + h sqrt(mu0 eps0 rp_nb), C = dt - h sqrt(mu0 eps0
// rp_face).  Planes of ny*nz entries, (k, j) order: out0 from rpx0 (face
// index 0), out1 from rpxEnd (face index 1).
void mur_x_coefs(int ny, int nz, float dt, float hx, float mu0, float eps0, const float *rpx0, const float *rpxEnd,
                 float4 *out0, float4 *out1);

// Which Mur x faces the fused kernel takes over.  The x-high faces need the
// cell at i = nx-2 from the face's wave: E^{n+1}_z from the previous lane
// through a wave shuffle, and E^n_y, E^n_z from the LDS column that lane
// staged, read after the plane's last barrier (only the staging wave's own
// program order keeps the next plane from overwriting it).  When nx-1 is a
// multiple of the cells one wave covers (tile_cells: 64, or 128 with two
// cells per lane; a 128-wide tile of one cell per lane is two waves of 64)
// the face is the first cell of a wave, that neighbour belongs to another
// wave or tile, and both x-high faces stay in the Mur kernel.
// The fused kernel stages the coefficients of its z-chunk in LDS, so chunks
// longer than kFusedMurMaxChunk planes leave the x faces to the Mur kernel.
constexpr int kFusedMurMaxChunk = 16;
inline int fused_mur_x(int nx, int kchunk, int tile_cells)
{
    if (kchunk > kFusedMurMaxChunk) return 0;
    return (nx - 1) % tile_cells == 0 ? 2 : 1;
}
// Mur segments (in mur_segments order) left to the Mur kernel.
inline unsigned mur_skip_mask(int mur_x)
{
    // segments 4, 5: Ey at i = 0, nx-1; 8, 9: Ez at i = 0, nx-1
    return mur_x == 1 ? (1u << 4 | 1u << 5 | 1u << 8 | 1u << 9) : mur_x == 2 ? (1u << 4 | 1u << 8) : 0u;
}

// Row pitch (elements) of the fused layout: rows start on 128-byte lines, so
// the 256-byte row segment a wave64 of fp32 lanes reads (or a 128-byte one
// of 16-bit lanes) never straddles an extra line.  Measured on MI355X: a
// 100x100x130 grid streams at 2.5 TB/s with rows packed (pitch = nx) and a
// 128^3 one at 4.4 TB/s.  Rows of 4 KiB or more whose length is a multiple
// of 2 KiB get one more line: such strides map the rows of a tile, and the
// planes of a z-march, onto the same DRAM channels.  Fused kernel, box
// phantom, best z-chunk each way (Gcell-updates/s): 2048x512x300 55.9 ->
// 63.1, 1536x1024x200 55.6 -> 62.7, 2048x1024x150 52.1 -> 60.8,
// 1000x1000x300 (pitch 1024) 56.4 -> 59.9, 1024x1024x200 60.2 -> 61.7; 2 KiB
// rows (512 fp32) gain nothing (512^3 61.5 -> 61.7), 1 KiB rows lose.
inline int fused_pitch(int nx, size_t esz)
{
    const int q = (int)(128 / esz);
    const int p = (nx + q - 1) / q * q;
    const size_t bytes = (size_t)p * esz;
    return bytes >= 4096 && bytes % 2048 == 0 ? p + q : p;
}

// Guard (in elements) the fused kernel needs around every field array and
// the id array so that its halo loads need no bounds checks: one plane, ten
// rows and 256 elements (tiles are at most 8 rows high and 128 cells wide).
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

// Electromagnetic energy sum_cells (E.D / 2 + mu0 |H|^2 / 2) of the newest
// levels, accumulated in fp64: per-block partial sums, then one block adds
// them in a fixed order (the result does not depend on scheduling).
struct EnergyArgs {
    const void *E[3], *D[3], *H[3];
    int nx, ny, nz, pitch;
    long long plane;
    double ed_factor, hh_factor;         // stored products -> SI: E.D and mu0 H^2
    double *partial;                     // kEnergyBlocks doubles
    double *out;                         // the sum is written here
};
constexpr int kEnergyBlocks = 1024;
hipError_t launch_energy(const EnergyArgs &a, int storage, hipStream_t s);

// Streaming probe (fdtd_stream_probe): 15 read and 9 written arrays of n floats.
struct StreamMix {
    const float *r[15];
    float *w[9];
    long long n;
    // row > 0: the 24 arrays interleaved by rows of `row` elements (row
    // records: the 24 rows of cells [q*row, (q+1)*row) contiguous, 15 read
    // then 9 written) from r[0]; r and w unused otherwise
    int row;
};
hipError_t launch_stream_mix(const StreamMix &p, int blocks, hipStream_t s);
// Blocks of the streaming probe resident on the whole device at once.
int stream_mix_resident_blocks();

// t0, t1 (optional): events that take the launch's start and end times.
hipError_t launch_fused_step(const StepArgs &a, int ty, int storage, hipStream_t s, hipEvent_t t0 = nullptr,
                             hipEvent_t t1 = nullptr);

// Two time steps in one launch (fp32 storage, the exact form).  A holds the
// roles of step n -> n+1 (level n+1 goes to A.exT.., A.dxT..); step n+1 ->
// n+2 writes level n+2 into the spare slot exS.., dxS.. and H^{n+3/2} into
// A.hxn..  A.tiles_* and A.kchunk describe the two-step tile grid
// (two_step_tiles); A.mur_x is not read (all six faces of step n+1 are
// applied in the kernel, those of step n+2 are left to the Mur kernel).
struct Step2Args {
    StepArgs A;
    void *exS, *eyS, *ezS;                   // E^{n+2}
    void *dxS, *dyS, *dzS;                   // D^{n+2}
    float jz2;                               // the source value of step n+1 -> n+2
    const float *rpy0, *rpyEnd, *rpz0, *rpzEnd;   // Mur planes of the y and z faces (MurArgs layout)
    float dt, hy, hz, mu0, eps0;
};
// Rows of the two-step kernel's 64-lane tiles; each tile stores the inner
// 62 x (rows - 2) cells.
constexpr int kTwoStepRows = 8;
constexpr int kTwoStepCols = 62;
inline void two_step_tiles(int nx, int ny, int nz, int kchunk, int rows, int *tx, int *ty, int *tz)
{
    *tx = (nx + kTwoStepCols - 1) / kTwoStepCols;
    *ty = (ny + rows - 3) / (rows - 2);
    *tz = (nz + kchunk - 1) / kchunk;
}
hipError_t launch_fused_step2(const Step2Args &a, int rows, hipStream_t s, hipEvent_t t0 = nullptr,
                              hipEvent_t t1 = nullptr);
hipError_t launch_mur(const MurArgs &a, int storage, hipStream_t s);
hipError_t launch_probes(const ProbeArgs &a, int storage, hipStream_t s);
// Halo planes of a z-slab: n runs of `words` 16-byte words each.
constexpr int kHaloFields = 5;
struct PlaneCopy {
    const uint4 *src[kHaloFields];
    uint4 *dst[kHaloFields];
    int n;
    long long words;
};
hipError_t launch_plane_copy(const PlaneCopy &p, hipStream_t s);
// The nine newest fields of a context as packed fp32 volumes in the
// reference's layout (field f at out + f*nx*ny*nz, OFFSET order), stored
// values times factor[f] in double (scales undone), as get_field computes
// them on the host; fp32 storage copies the bits.
struct SnapArgs {
    const void *field[9];
    double factor[9];
    float *out;
    int nx, ny, nz, pitch;
    long long plane;
};
hipError_t launch_snapshot(const SnapArgs &a, int storage, hipStream_t s);
// Mur faces and one probe row in one launch (p.nprobe = 0: faces only).  The
// probe fields must be the newest levels, with E the Mur target a.exT..ezT.
hipError_t launch_mur_probes(const MurArgs &a, const ProbeArgs &p, int storage, hipStream_t s);
// Prefix sums of the 12 Mur face segments; segments whose bit is set in
// `skip` get no cells.
void mur_segments(int nx, int ny, int nz, int *seg_start, unsigned skip = 0);

// ---- batches of independent grids -------------------------------------
//
// A batch advances many grids with one persistent fused launch per storage
// class and step, one Mur launch and one probe launch.  The launches read
// everything that changes from device memory (the instance descriptors, the
// per-class work tables and a batch step counter), so a captured sequence of
// steps replays unchanged while grids join and leave between replays.

constexpr int kBatchMaxInst = 512;        // descriptors per batch
constexpr int kBatchMaxProbes = 64;       // probes per instance
// Kernel classes of a batch: the four storage forms with one cell per lane
// (class = Storage value), then fp16 and bf16 with two cells per lane.
constexpr int kBatchClasses = 6;
inline int batch_class(int storage, int cpl)
{
    return cpl == 2 ? (storage == kStoreF16 ? 4 : 5) : storage;
}
inline int class_storage(int cls) { return cls < 4 ? cls : (cls == 4 ? kStoreF16 : kStoreBF16); }

// Cells per lane of the fused kernel for a grid nx wide: 16-bit storage runs
// two per lane (128-cell tiles) unless that leaves clearly more lanes idle
// than 64-cell tiles (a pair tile costs ~0.82 of a single one per cell, as
// measured at 256^3); `requested` 1 or 2 overrides.
inline int auto_lane_cells(int nx, size_t esz, int requested)
{
    if (esz != 2) return 1;
    if (requested == 1 || requested == 2) return requested;
    const double e1 = (double)nx / (64.0 * ((nx + 63) / 64)), e2 = (double)nx / (128.0 * ((nx + 127) / 128));
    return e2 >= 0.85 * e1 ? 2 : 1;
}

// Default launch shape of a grid launched alone (tile rows ty given), from
// sweeps on MI355X of 9 grid sizes (64^3 to 320x320x880) x tile rows {2, 4,
// 8} x z-chunks {4, 8, 16}, fp32, and of 8 sizes x {1, 2} cells per lane,
// fp16 and bf16:
//  * pairs (128-cell tiles, one 32-bit access per stream) beat single cells
//    where 128-cell rows are nearly full and the launch fills the GPU (fp16
//    256^3: 203 against 267 us; 128^3: 35 against 45), and lose elsewhere
//    (320 wide: 1115 against 1101; 200 wide: 53 against 48), also where
//    64-cell tiles would waste as many lanes;
//  * the z-chunk trades prologues (each block reloads and recomputes the
//    plane below its chunk) against the partial last round of blocks.  A
//    launch of B blocks on R resident ones costs about ceil(B / R) rounds
//    of kchunk + 2 planes; the chunk of least cost among 4, 8 and 16 is the
//    measured best, or within 2%, at every size but 200x100x100 (within 6%):
//    fp32 320x320x880 16 planes, 1544 against 1620 us at 4; 192^3 8 planes,
//    138 against 148 (16: 153); fp16 320x320x880 16 planes, 1101 against 1251.
// Resident blocks per CU: 4 waves per SIMD for pairs (their VGPR budget), 5
// for single cells, over ty waves a block.
inline long long alone_blocks(int nx, int ny, int nz, int width, int ty, int kchunk)
{
    return (long long)((nx + width - 1) / width) * ((ny + ty - 1) / ty) * ((nz + kchunk - 1) / kchunk);
}
inline long long alone_resident(int cus, int ty, int cpl)
{
    return (long long)cus * std::max(1, 4 * (cpl == 2 ? 4 : 5) / ty);
}
inline int auto_lane_cells_alone(int nx, int ny, int nz, int ty, int kchunk, size_t esz, int requested, int cus)
{
    if (esz != 2) return 1;
    if (requested == 1 || requested == 2) return requested;
    if ((double)nx / (128.0 * ((nx + 127) / 128)) < 0.9) return 1;
    return alone_blocks(nx, ny, nz, 128, ty, kchunk > 0 ? kchunk : 4) >= alone_resident(cus, ty, 2) ? 2 : 1;
}
inline int auto_kchunk_alone(int nx, int ny, int nz, int ty, int cpl, int cus)
{
    const long long R = alone_resident(cus, ty, cpl);
    int best = 4;
    long long best_cost = -1;
    for (int kc : {4, 8, 16}) {
        const long long rounds = (alone_blocks(nx, ny, nz, 64 * cpl, ty, kc) + R - 1) / R;
        const long long cost = rounds * (kc + 2);
        if (best_cost < 0 || cost <= best_cost) { best = kc; best_cost = cost; }
    }
    return std::min(best, nz);
}
constexpr int kMurBlock = 256;            // Mur face cells per block

// Device descriptor of one grid of a batch.  Instance step n (0-based) runs
// at batch step start + n for n < nsteps; outside that window the instance
// is left untouched.
struct BatchInst {
    void *E[3][3];                 // [ring slot][component]
    void *D[3][3];
    void *H[2][3];                 // [buffer][component]; buffer n % 2 holds H^{n-1/2}
    const uint8_t *ids;            // with fused_guard() bytes before and after
    const MatCoef *lut;
    const float *jz;               // per-step source values in storage units
    const float *rp[6];            // rpx0 rpy0 rpz0 rpxEnd rpyEnd rpzEnd
    const int *probes;             // nprobe x (field, i, j, k)
    double *probe_out;             // [step][probe], nsteps rows
    double factor[9];              // stored value -> SI, per field
    long long plane;
    int nmat, jz_len, jz_hold3, start, nsteps, storage;
    int nx, ny, nz, pitch;
    unsigned guard;
    float ch[3], cd[3];
    int kchunk, tiles_x, tiles_y, tiles_z;
    int mur_x;                     // see StepArgs
    const float4 *murx0, *murx1;   // see StepArgs
    int cpl;                       // cells per lane (kernel class)
    int zskip;                     // see StepArgs
    int nprobe, probe_rows;        // probe_out holds probe_rows rows
    float dt, hx, hy, hz, mu0, eps0;
    int mur_seg[13];
};

// Work list of one storage class: instance indices and prefix sums of their
// fused tiles.
struct BatchClass {
    int n, total;
    int idx[kBatchMaxInst];
    int start[kBatchMaxInst + 1];
};

// Work lists shared by all classes: Mur blocks per instance (prefix sums in
// units of kMurBlock face cells) and the flattened probe records.
struct BatchShared {
    int step;                      // batch step counter (advanced on the device)
    int nmur, mur_total;
    int mur_idx[kBatchMaxInst];
    int mur_start[kBatchMaxInst + 1];
    int nrec;
    int rec[kBatchMaxInst * kBatchMaxProbes][2];   // (instance, probe)
};

struct BatchLaunch {
    const BatchInst *inst;
    const BatchClass *cls;         // the launch's class
    BatchShared *sh;
    int off;                       // step offset from sh->step
    int xcd_runs;                  // tiles per XCD run (kBatchXcdRun), 0 = identity order
};
// Tile order of the batch's persistent kernel: each group of 8 runs of this
// many tiles gives one run to each XCD, so consecutive tiles (rows along x
// and their neighbouring rows, which read the same halo lines) share an L2.
// Batched set on MI355X (64 grids, fp32): 43.06 against 42.75 Gcell-updates/s
// for the identity order; runs of 4 and 16 tiles gain less, an eighth of each
// round of resident blocks per XCD loses 2 %.
constexpr int kBatchXcdRun = 8;

// LDS bytes of the batch's fused kernel (coefficient table of kMaxMaterials).
size_t batch_fused_lds();
hipError_t launch_fused_batch(const BatchLaunch &l, int ty, int cls, int grid, hipStream_t s);
hipError_t launch_mur_batch(const BatchLaunch &l, int grid, hipStream_t s);
hipError_t launch_batch_advance(BatchShared *sh, int steps, hipStream_t s);
// Resident blocks of the fused batch kernel of a class per CU.
int batch_fused_occupancy(int ty, int cls);

// ---- shared host helpers (fdtd_common.cpp) ------------------------------

struct GridGeom {
    float dt, dx, dy, dz, mu0, eps0;
};

int fdtd_fail(int code, const char *fmt, ...);
int fdtd_hip_fail(hipError_t e, const char *what);
uint16_t f32_to_f16(float f);
float f16_to_f32(uint16_t h);
uint16_t f32_to_bf16(float f);
float bf16_to_f32(uint16_t h);
void fdtd_scaled_coef(MatCoef &m, double dt, double eps0, double sg, double es, double ei, double tau, int debye_mode,
                      double ratio);
void fdtd_build_lut(const float *table, int nmat, float dt, float eps0, int debye_mode, int storage, double sE,
                    double sD, MatCoef *lut);
void fdtd_curl_coefs(const GridGeom &g, int storage, double sE, double sD, double sH, float ch[3], float cd[3]);
float fdtd_jz_stored(float jz, int storage, float eps0, double sD);
float fdtd_jz_at(const float *jz, int jz_len, int jz_limit, int jz_mode, int n);
void fdtd_auto_scales(const float *jz, int jz_len, float eps0, const FdtdExt &e, double *sE, double *sD, double *sH);
double fdtd_field_factor(int storage, float mu0, float eps0, double sE, double sD, double sH, int field);
size_t fdtd_storage_bytes(int storage);
void fdtd_to_storage(int storage, const float *in, size_t n, double factor, void *out);
void fdtd_from_storage(int storage, const void *in, size_t n, double factor, float *out);
// Row-pitched host images of a field (rows of nx values `pitch` apart,
// padding zeroed) and of an id array, for one contiguous device copy.
void fdtd_to_storage_pitched(int storage, const float *in, int nx, size_t rows, int pitch, double factor, void *out);
void fdtd_from_storage_pitched(int storage, const void *in, int nx, size_t rows, int pitch, double factor, float *out);
void fdtd_pitch_bytes(const unsigned char *in, int nx, size_t rows, int pitch, unsigned char *out);
int fdtd_check_ext(const FdtdExt &e);

// Phantom input (fdtd_phantom.cpp).  phantom_read_table fills a 95-row
// {sigma, eps_s, eps_inf, tau} table and returns the rows read;
// phantom_read_slice fills nx*ny ids of one PGM slice and returns the
// largest code; both return a negative FDTD_E* code with a message in err.
#include <string>
#include <vector>
int phantom_read_table(const char *path, float *specs, std::string &err);
int phantom_read_slice(const char *path, int nx, int ny, int mode, unsigned char *ids, std::vector<char> &buf,
                       std::string &err);
int fdtd_check_sources(int nx, int ny, int nz, int nsrc, const int *xyz);

// Device state of one grid on the drop-in path (the reference's FdtdField
// arrays, fdtd.h:37-54) and its probe buffers; shared by fdtd_dropin_solve
// (which allocates it) and fdtd_dropin_solve_noalloc (which carves it from a
// static device array).
struct DropinState {
    float *E[3][3], *D[3][3], *H[3];
    float *mat[4];                 // sigma, eps_s, eps_inf, tau per cell
    float *rp[6];
    int *probes = nullptr;         // nprobe x (field, i, j, k)
    double *probe_out = nullptr;   // steps x nprobe
};
// Argument checks of the drop-in solve; fills the effective FdtdExt.
int dropin_check(const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *in, const double *probes_host,
                 FdtdExt *out);
// Uploads the symbols, runs inst->steps steps of the five entry points on
// `stream` over state whose fields start at zero, copies the probe rows to
// probes_host and waits for the stream.
int dropin_run(const FdtdParams *p, const FdtdExt &e, const FdtdInstance *in, const DropinState &st,
               double *probes_host, hipStream_t stream);
int fdtd_check_probes(int nx, int ny, int nz, int nprobe, const int *probes);
// Largest |value| of each probe over `rows` rows of a series (SI units, row
// major, nprobe per row), in stored units (divided by its field's storage
// factor).  Returns how many probes saw a signal (peak > 0) that peaked
// below fp16's smallest normal number under fp16 storage (0 for the other
// storages).
int fdtd_probe_peaks(int storage, const double *series, int rows, int nprobe, const int *fields, const double *factor,
                     double *peak);

#define HIP_TRY(call)                                              \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return fdtd_hip_fail(e_, #call);     \
    } while (0)
