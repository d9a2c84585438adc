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
#define FDTD_ENOMEM   (-5)  /* does not fit the device memory budget or the slot table */

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
    int kchunk;         /* z planes per fused block: 0 = auto (4) */
    int xcd_remap;      /* block -> tile order: 0 identity, 1 one contiguous run per XCD, 2 adjacent
                           z-chunk layers on the 8 XCDs, 3 (default) adjacent rows of tiles along x on
                           the 8 XCDs, 4 each layer in 8 strips of rows, one per XCD */
    int storage;        /* FDTD_STORE_F32 (default), _F16, _BF16 or _F32_FAST */
    float scale_e;      /* 16-bit storage: E stored as E / scale_e (0 = auto) */
    float scale_d;      /* 16-bit storage: D stored as D / (eps0 scale_d) (0 = auto) */
    float scale_h;      /* 16-bit storage: H stored as eta0 H / scale_h (0 = auto) */
    int lane_cells;     /* 16-bit storage: cells per lane of the fused kernel (1 or 2; 0 = auto: 2 unless
                           128-cell tiles leave clearly more lanes idle than 64-cell ones) */
    int tile_x;         /* threads per tile row with one cell per lane: 0 = 64, or 64, 128 */
    int load_zero_terms; /* fp32 storage: 0 (default) skips loading E^{n-1} and D^{n-1} on the half rows
                            of a wave (32 cells) whose cells' Debye update multiplies them by exactly
                            zero (tau = 0: vacuum and non-dispersive media), with results unchanged bit
                            for bit; 1 always loads them.  Off after fdtd_ctx_set_field or
                            fdtd_ctx_load. */
    int time_steps;     /* fp32 storage: time steps per fused launch, 0 = auto (1) or 1, 2.  With 2, fdtd_run
                           advances pairs of steps in one launch where it can (stepped from zero, no halo
                           exchange, no H probes, no energy sum or slice on the first step of a pair),
                           bit for bit as one step at a time; it holds a fourth E and D slot (24 bytes per
                           cell more). */
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
/* Soft Jz sources (0-based cells, 1 <= x < nx, 1 <= y < ny, 0 <= z < nz:
 * the cells updateSources can address, fdtd_calculations.cu:198-211, whose
 * Hy(x-1) and Hx(y-1) reads need x, y >= 1) sharing one waveform
 * jz[jz_len].  A source on an upper x/y face or on the top plane sets Dz
 * there although the D update skips those cells, as the reference does. */
int fdtd_ctx_set_sources(fdtd_ctx *ctx, int nsrc, const int *xyz, int jz_len, const float *jz);
/* Point probes (field id, i, j, k); one row of nprobe doubles is recorded on
 * the device after every step, up to `capacity` rows. */
int fdtd_ctx_set_probes(fdtd_ctx *ctx, int nprobe, const int *probes, int capacity);
/* Capture z-plane `k` of all nine fields every `interval` steps into pinned
 * host memory on a side stream; up to `capacity` slices are kept. */
int fdtd_ctx_set_slices(fdtd_ctx *ctx, int k, int interval, int capacity);

/* Record the field energy sum(E.D/2 + mu0 |H|^2/2) over all cells, of the
 * newest levels, every `interval` steps (up to `capacity` values),
 * accumulated in fp64 on the device in a fixed order. */
int fdtd_ctx_set_energy(fdtd_ctx *ctx, int interval, int capacity);
/* Waits for the device; returns the number of values and fills steps[]
 * (the step count after which each was taken) and energy[] (the sum of the
 * cell energy densities, J/m^3; times dx*dy*dz for joules), up to
 * max_rows. */
int fdtd_energy_read(fdtd_ctx *ctx, int *steps, double *energy, int max_rows);

/* level 0 = newest (after the last step), 1 = the one before, 2 = oldest
 * ring slot.  H has one level. */
int fdtd_ctx_set_field(fdtd_ctx *ctx, int field, int level, const float *host);
int fdtd_ctx_get_field(fdtd_ctx *ctx, int field, int level, float *host);
/* Device address of a field (for zero-copy callers); elements are float,
 * or 16-bit scaled values under 16-bit storage (see fdtd_ctx_scales), laid
 * out as fdtd_ctx_layout says. */
int fdtd_ctx_field_ptr(fdtd_ctx *ctx, int field, int level, void **dev);
/* Device layout of every field: cell (i, j, k) is element
 * k*plane + j*pitch + i; pitch >= nx rounds rows up to 128-byte lines,
 * plane = pitch*ny; elem_bytes is 4 or 2. */
int fdtd_ctx_layout(fdtd_ctx *ctx, int *pitch, long long *plane, int *elem_bytes);

/* z-slab decomposition (no counterpart in the reference, which runs one
 * grid on one device).  A context may hold one z-slab of a larger grid plus
 * a ghost plane on each side that has a neighbouring slab: local plane 0
 * below, local plane nz-1 above.  Before every step each interface
 * exchanges one message each way, and then both slabs step; their real
 * planes then equal the whole grid's bit for bit (same storage form and
 * scales, sources and probes given to the slab that owns the cell, the
 * global grid's rp planes sliced to the slab).
 *   FDTD_HALO_UP:   packed by the lower slab from its last real plane
 *                   (E^n x, y, z and H^{n-1/2} x, y), unpacked by the upper
 *                   slab into its ghost plane 0;
 *   FDTD_HALO_DOWN: packed by the upper slab from its first real plane
 *                   (E^n x, y), unpacked by the lower slab into its ghost
 *                   plane nz-1.
 * A message is a contiguous device buffer of fdtd_ctx_halo_bytes bytes
 * (whole pitched planes, one per field), 16-byte aligned: what a
 * point-to-point send/recv between two GPUs carries.  Pack and unpack are
 * one kernel launch each on `stream`. */
#define FDTD_HALO_UP 0
#define FDTD_HALO_DOWN 1
int fdtd_ctx_halo_bytes(fdtd_ctx *ctx, int dir, size_t *bytes);
int fdtd_ctx_halo_pack(fdtd_ctx *ctx, int dir, void *dst, hipStream_t stream);
/* src must hold a neighbour's fdtd_ctx_halo_pack message of the same step (the
 * zero-term rows of load_zero_terms = 0 rely on its cells having been stepped
 * by a fused kernel). */
int fdtd_ctx_halo_unpack(fdtd_ctx *ctx, int dir, const void *src, hipStream_t stream);
/* Storage scales in use (1 for FDTD_STORE_F32).  Auto scales, fixed at the
 * first step, come from the source waveform: E and D/eps0 take a third of
 * the accumulated kick max_t |sum jz| / eps0 over 4096, eta0 H 1.25 times
 * the largest kick max |jz| / eps0 over 4096. */
int fdtd_ctx_scales(fdtd_ctx *ctx, double *scale_e, double *scale_d, double *scale_h);

/* Checkpoint and restart: write every field level, the step count and the
 * storage scales to a binary file, bit for bit; load restores them into a
 * context of the same grid and storage.  Materials, sources, rp planes,
 * probes and slices are inputs and are set on the restoring context as on
 * the original. */
int fdtd_ctx_save(fdtd_ctx *ctx, const char *path);
int fdtd_ctx_load(fdtd_ctx *ctx, const char *path);

/* One full time step (H, D, sources, Debye E, Mur, probes) on `stream`:
 * the five launches of one step of the reference's loop (fdtd.cu:129-214,
 * updateHField ... updateMurBoundary, with the ring rotation of
 * fdtd.cu:115-530) as two launches, the fused H/D/source/E kernel and the
 * Mur faces with the step's probe row. */
int fdtd_step_fused(fdtd_ctx *ctx, hipStream_t stream);
/* `steps` steps; enqueues only (no host synchronisation). */
int fdtd_run(fdtd_ctx *ctx, int steps, hipStream_t stream);
int fdtd_ctx_steps_done(fdtd_ctx *ctx);

/* Enqueues on `stream` a copy of the newest level of all nine fields into
 * dev_out (device memory of 9*nx*ny*nz floats): field f (FDTD_DX ..
 * FDTD_EZ) at dev_out + f*nx*ny*nz, cell (i, j, k) at k*ny*nx + j*nx + i
 * (the reference's OFFSET layout, utils.h:15), in SI units whatever the
 * storage (the values fdtd_ctx_get_field returns).  Stepping may continue
 * on `stream` at once: the copy is ordered before the next step.  The
 * device half of an overlapped result dump (the D2H copy and the text
 * writers of fdtd.cu:134-155, 1136-1194 then run beside the steps). */
int fdtd_ctx_snapshot(fdtd_ctx *ctx, float *dev_out, hipStream_t stream);

/* Copies up to max_rows recorded probe rows to host (synchronises the
 * context's streams); returns the number of rows or a negative code. */
int fdtd_probe_read(fdtd_ctx *ctx, double *host, int max_rows);
/* Precision check for 16-bit storage (no counterpart in the reference, whose
 * fields are fp32).  peak[p] = the largest |value| probe p has recorded so
 * far, in stored units (the SI value over its field's scale; SI for fp32
 * storage); nprobe must equal the context's probe count.  Returns how many
 * probes have seen a signal that peaked below 2^-14, fp16's smallest normal
 * number, under fp16 storage (0 for the others), or a negative code.  Such a probe's field
 * never had fp16's 11 significant bits: one scale per field cannot hold
 * both a strong source's near field and what it observes there, and its
 * error against FP64 can reach O(1) (fp32 storage is the remedy). */
int fdtd_probe_peak_units(fdtd_ctx *ctx, double *peak, int nprobe);
/* Waits for the slice copies; returns the number captured and fills
 * steps[] (the step count after which each slice was taken) and data
 * (slices x 9 fields x nx*ny floats) up to max_slices. */
int fdtd_slices_read(fdtd_ctx *ctx, int *steps, float *data, int max_slices);

/* Autotuner: times every launch shape (tile rows {2,4,8} x z planes per
 * block {3,4,6,8,16} x tile rows of 64 or 128 threads, and two cells per lane
 * under 16-bit storage) on this
 * grid for trial_steps steps each, keeps the fastest and returns the
 * context to its zero state at step 0 (probe rows and captures restart).
 * Needs a context that has not stepped and whose fields were not set. */
int fdtd_ctx_autotune(fdtd_ctx *ctx, int trial_steps, hipStream_t stream, int *tile_y, int *kchunk, int *lane_cells);

/* The launch shape chosen for the context's fused kernel. */
int fdtd_ctx_launch_shape(fdtd_ctx *ctx, int *tile_x, int *tile_y, int *kchunk, int *blocks);
/* The launch shape fdtd_ctx_create would take for this grid and extension
 * (no counterpart in the reference, whose launch geometry is fixed at
 * (ceil(nx/128), ny, nz) x 128 threads, fdtd.cu:129-490): 4 tile rows; for
 * 16-bit storage two cells per lane (tile_x 128) where 128-cell rows are at
 * least 90% full and a pair launch fills the GPU; the z-chunk of least cost
 * among 4, 8 and 16 planes, a launch of B blocks on R resident ones costing
 * ceil(B/R) rounds of kchunk + 2 planes.  ext's tile_y, kchunk, lane_cells
 * and tile_x fix their parts.  cus: compute units of the target GPU (0 = the
 * current device's, or 256 when none can be queried).  Allocates nothing and
 * launches nothing. */
int fdtd_default_launch_shape(const FdtdParams *p, const FdtdExt *ext, int cus, int *tile_x, int *tile_y,
                              int *kchunk, int *lane_cells);

/* Times every `every`-th fused-kernel launch (the first, then each `every`
 * launches; every >= 1), up to `capacity` of them (capacity 0 disables):
 * a HIP event pair bound to the launch's dispatch takes the kernel's own
 * start and end.  Timing every launch costs the step sequence about 1.5%,
 * so sampling keeps the timed run's wall clock close to an untimed one.
 * fdtd_ctx_kernel_time waits for the recorded pairs and returns their
 * summed duration.  Only launches of one kind are timed, the kind of the
 * first timed launch: fdtd_ctx_timed_steps returns its steps per launch (1,
 * or 2 for two-step launches; 0 before any launch was timed). */
int fdtd_ctx_set_timing(fdtd_ctx *ctx, int capacity, int every);
int fdtd_ctx_kernel_time(fdtd_ctx *ctx, double *ms_total, int *launches);
int fdtd_ctx_timed_steps(fdtd_ctx *ctx);

/* Memory-ceiling calibration (no counterpart in the reference): streams the
 * fused step's access mix without its arithmetic, 15 arrays read and 9
 * written element for element (cells floats each, 64 Ki elements apart as
 * in a context, every access nontemporal), `reps` timed passes on `blocks`
 * blocks of 256 threads (0 = as many as fit resident).  Allocates and frees
 * 96 bytes per cell.  gbs receives 96 * cells bytes over the mean pass
 * time, ms_per_pass (optional) that time. */
int fdtd_stream_probe(long long cells, int blocks, int reps, double *gbs, double *ms_per_pass);
/* The same mix in another layout: row = 0 is fdtd_stream_probe's 24
 * separate arrays; row > 0 interleaves them by rows of `row` elements (the
 * 24 rows that cells [q*row, (q+1)*row) touch are contiguous, a 96*row-byte
 * record), so every wave's accesses fall in one record.  cells must be a
 * multiple of row. */
int fdtd_stream_probe_layout(long long cells, int row, int blocks, int reps, double *gbs, double *ms_per_pass);

/* ---- voxel phantoms (fdtd.cu:875-939) ------------------------------------ */

#define FDTD_PGM_REFERENCE 0  /* nx*ny codes in file order whatever the image width (fdtd.cu:920-933) */
#define FDTD_PGM_IMAGE     1  /* the code of image pixel (ix, iy); cells outside the image are air (0) */

/* The tissue table as the reference reads it (fdtd.cu:887-904): table95
 * receives 95 rows {sigma, eps_s, eps_inf, tau_d}, row = tissue code.
 * Returns the rows read. */
int fdtd_load_tissue_table(const char *specs_path, float *table95);
/* Tissue codes of slices <dir>/v1_%05d.pgm (k+1 for plane k) into ids
 * (nx*ny*nz bytes, x fastest; id = tissue code, 0..94), parsed on `threads`
 * threads (0 = all cores).  Use with fdtd_ctx_set_materials(ctx, ids, 95,
 * table95).  Returns the largest code seen. */
int fdtd_load_phantom_ids(const char *dir, int nx, int ny, int nz, int mode, int threads, unsigned char *ids);

/* ---- batches of independent grids --------------------------------------
 *
 * A batch holds many grids of different sizes, materials, sources, storage
 * forms and step counts in one device arena.  Each step of the batch is one
 * persistent fused kernel per storage form present, one Mur launch and one
 * probe launch over every running grid, replayed from captured HIP graphs;
 * grids join (fdtd_batch_add) and leave (fdtd_batch_remove) between calls.
 * A grid's results are those of a single fdtd_ctx with the same inputs, bit
 * for bit (the per-cell arithmetic does not depend on the batch). */

typedef struct fdtd_batch fdtd_batch;

#define FDTD_BATCH_MAX_INSTANCES 512
#define FDTD_BATCH_MAX_PROBES 64

typedef struct
{
    const unsigned char *ids;   /* nx*ny*nz material ids (x fastest), or NULL: all material 0 */
    int nmat;                   /* rows of table, 1..128 */
    const float *table;         /* nmat x {sigma, eps_s, eps_inf, tau_d} */
    const float *rp[6];         /* Mur planes rpx0 rpy0 rpz0 rpxEnd rpyEnd rpzEnd; NULL: all 1 */
    int nsrc;
    const int *src_xyz;         /* nsrc x (x, y, z), 0-based, as fdtd_ctx_set_sources */
    int jz_len;
    const float *jz;            /* source waveform */
    int steps;                  /* time steps to run, >= 1 */
    int nprobe;                 /* 0..FDTD_BATCH_MAX_PROBES */
    const int *probes;          /* nprobe x (field, i, j, k); one row per step */
} FdtdInstance;

typedef struct
{
    int instances;              /* added and not removed */
    int running;                /* of them, not yet at their step count */
    long long step;             /* batch steps run */
    size_t arena_bytes;         /* device arena (the memory budget) */
    size_t used_bytes;
    size_t largest_free;        /* largest region an instance can still get */
    int tile_y, kchunk;         /* fused launch shape */
    int graph_steps;            /* steps per captured graph */
    int grid[4];                /* persistent blocks per fused launch, per storage form (largest) */
    int tiles[4];               /* fused tiles per step of the running grids, per storage form */
    long long cells[4];         /* cells of the running grids, per storage form */
} FdtdBatchInfo;

/* hbm_bytes: size of the device arena that holds every grid's state. */
int fdtd_batch_create(size_t hbm_bytes, fdtd_batch **out);
int fdtd_batch_destroy(fdtd_batch *b);
/* Arena bytes fdtd_batch_add would take for this grid. */
int fdtd_batch_instance_bytes(const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *inst, size_t *bytes);
/* Uploads a grid (zero initial fields, as copyDataToDevice does) and gives
 * it an id; it takes its first step in the next fdtd_batch_run.  Of ext,
 * debye_mode, jz_index_mode, jz_limit, storage and the scales apply; the
 * launch shape is the batch's.  FDTD_ENOMEM when the arena or the slot
 * table is full. */
int fdtd_batch_add(fdtd_batch *b, const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *inst, int *id);
/* Waits for the batch's work and frees the grid's arena region. */
int fdtd_batch_remove(fdtd_batch *b, int id);
/* Runs `steps` batch steps on `stream` (enqueue only); a grid stops at its
 * own step count. */
int fdtd_batch_run(fdtd_batch *b, int steps, hipStream_t stream);
int fdtd_batch_status(fdtd_batch *b, int id, int *steps_done, int *steps_total);
/* Waits for the device; copies up to max_rows probe rows; returns rows. */
int fdtd_batch_probe_read(fdtd_batch *b, int id, double *host, int max_rows);
/* fdtd_probe_peak_units for grid `id` of a batch. */
int fdtd_batch_probe_peak_units(fdtd_batch *b, int id, double *peak, int nprobe);
/* Waits for the device; level as fdtd_ctx_get_field, after the grid's
 * steps so far. */
int fdtd_batch_get_field(fdtd_batch *b, int id, int field, int level, float *host);
/* tile_y in {2,4,8}, kchunk >= 1, graph_steps >= 1 (0 keeps a value). */
int fdtd_batch_set_shape(fdtd_batch *b, int tile_y, int kchunk, int graph_steps);
/* Times every (tile_y in {2,4,8}, kchunk in {3,4,6,8,16}) candidate on the batch's own grids for
 * trial_steps steps each, keeps the fastest and restores the grids to
 * their zero initial state.  Every grid must be unstepped. */
int fdtd_batch_autotune(fdtd_batch *b, int trial_steps, hipStream_t stream, int *tile_y, int *kchunk);
int fdtd_batch_info(fdtd_batch *b, FdtdBatchInfo *info);

/* One grid solved through the five drop-in entry points of
 * fdtd_calculations.h, as the reference's main loop drives them
 * (fdtd.cu:115-530): per-voxel material arrays, the reference's limits (16
 * sources, the first 512 jz values), fp32 storage.  Allocates, steps
 * inst->steps times on `stream`, writes nprobe x steps probe rows to
 * probes_host and frees.  Uses the process-global symbols of
 * copySymbolsToDevice: calls must not overlap.  The sequential baseline of
 * batched and served runs. */
int fdtd_dropin_solve(const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *inst, double *probes_host,
                      hipStream_t stream);

/* fdtd_dropin_solve without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops and harnesses that forbid hipMalloc).  The
 * grid's state (21 field and 4 material arrays, rp planes, ids, table,
 * probe rows) lives in a static device array of 64 MiB per GPU, held until
 * the process ends: a grid whose state exceeds it (about 640 K cells, e.g.
 * 86^3) is refused with FDTD_ERANGE.  Same inputs, outputs and results as
 * fdtd_dropin_solve, bit for bit; calls no hipMalloc/hipFree.  All calls on
 * one GPU share that array (and the process-global symbols), so calls that
 * overlap on one GPU race; each call clears the state it uses first. */
int fdtd_dropin_solve_noalloc(const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *inst,
                              double *probes_host, hipStream_t stream);

/* ---- serving ------------------------------------------------------------
 *
 * A server owns a batch and a worker thread.  Requests (a grid, its inputs
 * and probes) queue on submit; between graph replays the worker admits
 * queued requests in arrival order while the slot limit and the memory
 * budget allow (a request that does not fit lets later ones that fit pass),
 * steps every admitted grid together, and retires finished ones: their
 * probe series are copied out and their memory returns to the budget. */

typedef struct fdtd_serve fdtd_serve;

typedef struct
{
    size_t hbm_bytes;           /* memory budget (the batch arena) */
    int max_slots;              /* grids on the device at once, <= FDTD_BATCH_MAX_INSTANCES */
    int segment_steps;          /* steps per replay between admissions (0 = 8) */
    int tile_y, kchunk;         /* fused launch shape (0 = defaults) */
} FdtdServeConfig;

#define FDTD_REQ_QUEUED  0
#define FDTD_REQ_RUNNING 1
#define FDTD_REQ_DONE    2
#define FDTD_REQ_FAILED  3

typedef struct
{
    int state;                  /* FDTD_REQ_* */
    int error;                  /* FDTD_E* or hipError_t when FAILED */
    int rows, nprobe;           /* probe series shape */
    double t_submit, t_admit, t_done;   /* seconds since fdtd_serve_create */
    long long cells;            /* nx*ny*nz */
    int steps;
} FdtdRequestInfo;

typedef struct
{
    int submitted, completed, failed, queued, running;
    double t_first_submit, t_last_done;   /* seconds since fdtd_serve_create */
    double solves_per_s;        /* completed / (t_last_done - t_first_submit) */
    double p50_latency_s, p99_latency_s, mean_latency_s;   /* t_done - t_submit of completed */
    double cell_updates;        /* sum of cells x steps of completed */
    long long segments;         /* graph replays */
    size_t peak_bytes;          /* largest arena use */
    int peak_running;
} FdtdServeStats;

int fdtd_serve_create(const FdtdServeConfig *cfg, fdtd_serve **out);
/* Stops the worker after the admitted grids finish; queued requests fail. */
int fdtd_serve_destroy(fdtd_serve *s);
/* Copies the inputs; thread-safe; returns at once. */
int fdtd_serve_submit(fdtd_serve *s, const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *inst,
                      long long *request);
/* 0 when done or failed, 1 on timeout (timeout_s < 0: wait without limit). */
int fdtd_serve_wait(fdtd_serve *s, long long request, double timeout_s);
/* Copies up to max_rows probe rows of a finished request; returns rows. */
int fdtd_serve_result(fdtd_serve *s, long long request, double *probes, int max_rows, FdtdRequestInfo *info);
/* Frees a finished request's host results. */
int fdtd_serve_release(fdtd_serve *s, long long request);
/* Waits until nothing is queued or running (1 on timeout). */
int fdtd_serve_drain(fdtd_serve *s, double timeout_s);
int fdtd_serve_stats(fdtd_serve *s, FdtdServeStats *st);

#ifdef __cplusplus
}
#endif

#endif
