/*
 * fdtd_amd.h -- context-based entry points of libfdtd_amd.so beside the
 * reference's drop-in set (fdtd_calculations.h).  No counterparts exist in
 * the reference; these carry the fused MI355X step and its extensions:
 * per-context state (re-entrant, no globals), uint8 material ids with a
 * per-material coefficient table, full-length source waveforms, device-side
 * probe series and z-slice capture.
 *
 * Return values: 0 on success, a negative FDTD_E* code for a rejected
 * argument, or a positive hipError_t value for a failed HIP call.
 * fdtd_last_error() returns a message for the last failure on the calling
 * thread.
 */
#ifndef FDTD_AMD_H
#define FDTD_AMD_H

#include <hip/hip_runtime_api.h>

#include "fdtd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FDTD_DEBYE_REF   0  /* d = 2b - sigma dt^2 (fdtd_calculations.cu:144) */
#define FDTD_DEBYE_FIXED 1  /* d = a + b - sigma dt^2 */

#define FDTD_JZ_HOLD3 0     /* step n uses jz[3*floor(n/3)] (fdtd.cu:172,310,448) */
#define FDTD_JZ_STEP  1     /* step n uses jz[n] */

#define FDTD_STORE_F32  0   /* fp32 fields: the reference's arithmetic, bit for bit */
#define FDTD_STORE_F16  1   /* fp16 fields, fp32 arithmetic, per-field scales */
#define FDTD_STORE_BF16 2   /* bf16 fields, fp32 arithmetic, per-field scales */
#define FDTD_STORE_F32_FAST 3 /* fp32 fields, the scaled-unit fp32 arithmetic of the 16-bit forms */

#define FDTD_EINVAL   (-1)  /* bad argument */
#define FDTD_ERANGE   (-2)  /* index or size out of range */
#define FDTD_ESTATE   (-3)  /* call out of order */
#define FDTD_EIO      (-4)  /* file could not be read */

/* Field ids, in the column order of the reference's probe writer
 * (fdtd.cu:1223). */
#define FDTD_DX 0
#define FDTD_DY 1
#define FDTD_DZ 2
#define FDTD_HX 3
#define FDTD_HY 4
#define FDTD_HZ 5
#define FDTD_EX 6
#define FDTD_EY 7
#define FDTD_EZ 8

typedef struct
{
    int debye_mode;     /* FDTD_DEBYE_FIXED (default) or FDTD_DEBYE_REF */
    int jz_index_mode;  /* FDTD_JZ_STEP (default) or FDTD_JZ_HOLD3 */
    int jz_limit;       /* jz entries used; 0 = the whole table, 512 = the reference's device table */
    int tile_y;         /* rows per fused tile: 0 = auto (4), else 2, 4 or 8 */
    int kchunk;         /* z planes per fused block: 0 = auto (8) */
    int xcd_remap;      /* block -> tile order: 0 identity (default), 1 one contiguous run
                           per XCD, 2 adjacent z-chunk layers on the 8 XCDs */
    int storage;        /* FDTD_STORE_F32 (default), _F16, _BF16 or _F32_FAST */
    float scale_e;      /* 16-bit storage: E stored as E / scale_e (0 = auto) */
    float scale_d;      /* 16-bit storage: D stored as D / (eps0 scale_d) (0 = auto) */
    float scale_h;      /* 16-bit storage: H stored as eta0 H / scale_h (0 = auto) */
} FdtdExt;

typedef struct fdtd_ctx fdtd_ctx;

const char *fdtd_last_error(void);

int fdtd_ext_defaults(FdtdExt *ext);

/* Allocates all device state of one grid (E and D rings of 3 slots,
 * ping-pong H, material ids) and zeroes it, like copyDataToDevice
 * (fdtd.cu:1085-1119).  Uses nx/ny/nz, dt, dx/dy/dz, mu0, eps0 of params. */
int fdtd_ctx_create(const FdtdParams *params, const FdtdExt *ext, fdtd_ctx **out);
int fdtd_ctx_destroy(fdtd_ctx *ctx);

/* ids: nx*ny*nz bytes, x fastest; id < nmat <= 128.  table: nmat rows of
 * {sigma, eps_s, eps_inf, tau_d}, the tissue-table columns
 * (fdtd.cu:895-900). */
int fdtd_ctx_set_materials(fdtd_ctx *ctx, const unsigned char *ids, int nmat, const float *table);
/* The reference's four per-voxel arrays (fdtd.h:50-51); at most 128
 * distinct tuples. */
int fdtd_ctx_set_material_arrays(fdtd_ctx *ctx, const float *sigma, const float *epsS,
                                 const float *epsI, const float *tauD);
/* The six Mur planes in the reference's layout (fdtd.cu:716-722). */
int fdtd_ctx_set_rp(fdtd_ctx *ctx, const float *rpx0, const float *rpy0, const float *rpz0,
                    const float *rpxEnd, const float *rpyEnd, const float *rpzEnd);
/* Soft Jz sources (0-based cells, 1 <= x < nx-1, 1 <= y < ny-1,
 * 0 <= z < nz-1) sharing one waveform jz[jz_len]. */
int fdtd_ctx_set_sources(fdtd_ctx *ctx, int nsrc, const int *xyz, int jz_len, const float *jz);
/* Point probes (field id, i, j, k); one row of nprobe doubles is recorded on
 * the device after every step, up to `capacity` rows. */
int fdtd_ctx_set_probes(fdtd_ctx *ctx, int nprobe, const int *probes, int capacity);
/* Capture z-plane `k` of all nine fields every `interval` steps into pinned
 * host memory on a side stream; up to `capacity` slices are kept. */
int fdtd_ctx_set_slices(fdtd_ctx *ctx, int k, int interval, int capacity);

/* level 0 = newest (after the last step), 1 = the one before, 2 = oldest
 * ring slot.  H has one level. */
int fdtd_ctx_set_field(fdtd_ctx *ctx, int field, int level, const float *host);
int fdtd_ctx_get_field(fdtd_ctx *ctx, int field, int level, float *host);
/* Device address of a field (for zero-copy callers); elements are float,
 * or 16-bit scaled values under 16-bit storage (see fdtd_ctx_scales). */
int fdtd_ctx_field_ptr(fdtd_ctx *ctx, int field, int level, void **dev);
/* Storage scales in use (1 for FDTD_STORE_F32).  Auto scales are 1/32 of
 * the largest per-step D/eps0 kick of the source waveform, fixed at the
 * first step. */
int fdtd_ctx_scales(fdtd_ctx *ctx, double *scale_e, double *scale_d, double *scale_h);

/* One full time step (H, D, sources, Debye E, Mur, probes) on `stream`. */
int fdtd_step_fused(fdtd_ctx *ctx, hipStream_t stream);
/* `steps` steps; enqueues only (no host synchronisation). */
int fdtd_run(fdtd_ctx *ctx, int steps, hipStream_t stream);
int fdtd_ctx_steps_done(fdtd_ctx *ctx);

/* Copies up to max_rows recorded probe rows to host (synchronises the
 * context's streams); returns the number of rows or a negative code. */
int fdtd_probe_read(fdtd_ctx *ctx, double *host, int max_rows);
/* Waits for the slice copies; returns the number captured and fills
 * steps[] (the step count after which each slice was taken) and data
 * (slices x 9 fields x nx*ny floats) up to max_slices. */
int fdtd_slices_read(fdtd_ctx *ctx, int *steps, float *data, int max_slices);

/* The launch shape chosen for the context's fused kernel. */
int fdtd_ctx_launch_shape(fdtd_ctx *ctx, int *tile_x, int *tile_y, int *kchunk, int *blocks);

/* Records a HIP event pair around each fused-kernel launch of the next
 * `capacity` steps, on the stream the kernel runs on (0 disables);
 * fdtd_ctx_kernel_time waits for them and returns the summed duration. */
int fdtd_ctx_set_timing(fdtd_ctx *ctx, int capacity);
int fdtd_ctx_kernel_time(fdtd_ctx *ctx, double *ms_total, int *launches);

#ifdef __cplusplus
}
#endif

#endif
