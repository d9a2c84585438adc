// Context-based API of libfdtd_amd (fdtd_amd.h): device state of one grid,
// the fused step sequence (fused H/D/E kernel -> Mur faces -> probes) and
// the overlapped z-slice capture on a side stream.
//
// Storage: fp32 fields reproduce the reference's arithmetic bit for bit.
// 16-bit fields (fp16 or bf16) hold E/sE, (D/eps0)/sD and (eta0 H)/sH; the
// per-field scales sE, sD, sH are folded into the step's coefficients, so
// the kernels never rescale a value, and are undone when results leave the
// device (probes, slices, get_field).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "fdtd_internal.h"

namespace {

// Extra elements between consecutive field arrays in the blob: staggers the
// arrays' starts across HBM channels (a no-compute 15-read/9-write stream
// probe measured ~4% more bandwidth with 64 Ki elements of stagger).
constexpr size_t kArraySkew = 65536;

}  // namespace

struct fdtd_ctx {
    int nx, ny, nz;
    size_t N;                            // cells
    size_t Np;                           // elements of one field array (plane * nz)
    int pitch;                           // elements between rows (fused_pitch)
    long long plane;                     // pitch * ny
    float dt, dx, dy, dz, mu0, eps0;
    FdtdExt ext;
    int storage;
    size_t esz;                          // bytes per stored element
    int ty, kchunk, tiles_x, tiles_y, tiles_z;
    int cpl = 1;                         // cells per lane of the fused kernel
    int txw = 64;                        // threads per tile row (one cell per lane)
    bool fields_set = false;             // set_field or load gave non-zero state

    long long guard = 0;                 // elements before/after the field arrays and the ids
    char *blob = nullptr;                // all field arrays
    unsigned char *ids_base = nullptr;   // allocation holding ids with its guards
    char *E[3][3], *D[3][3], *H[2][3];   // [slot][component], H[buffer][component]
    int hcur = 0;                        // H buffer holding the newest H
    // Two-step launches: a fourth E and D slot (allocated at the first one),
    // which takes level n+2 and is then swapped with the ring slot of n-1.
    char *spare_blob = nullptr;
    char *Es[3] = {nullptr, nullptr, nullptr}, *Ds[3] = {nullptr, nullptr, nullptr};
    int kchunk2 = 0;                     // z planes per two-step block
    bool halo_used = false;              // slab contexts step one step at a time
    unsigned char *ids = nullptr;
    MatCoef *lut = nullptr;
    int nmat = 0;
    std::vector<float> table;             // nmat x {sigma, eps_s, eps_inf, tau}
    bool lut_dirty = true;
    float *rp[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    float4 *murx = nullptr;              // x-face Mur coefficients: 2 planes of ny*nz
    std::vector<unsigned char> ids_host;  // material ids without source bits
    std::vector<int> src;                 // 3 per source
    std::vector<int> probe_field;         // field of each probe
    std::vector<float> jz;
    double sE = 1.0, sD = 1.0, sH = 1.0;  // 16-bit storage scales
    bool scales_set = false;

    int *probes = nullptr;
    double *probe_out = nullptr;
    int nprobe = 0, probe_cap = 0, probe_rows = 0;

    int slice_k = -1, slice_interval = 0, slice_cap = 0, slice_count = 0;
    char *slice_stage = nullptr;   // device: cap x 9 x plane elements
    char *slice_host = nullptr;    // pinned host: cap x 9 x plane elements
    std::vector<int> slice_steps;
    hipStream_t side = nullptr;
    hipEvent_t slice_ev = nullptr;

    int energy_interval = 0, energy_cap = 0, energy_rows = 0;
    double *energy_out = nullptr;        // device: energy_cap sums, then kEnergyBlocks partials
    std::vector<int> energy_steps;

    std::vector<hipEvent_t> tev;  // start/end pairs around fused launches
    int tcap = 0, tused = 0;
    int tevery = 1, tseen = 0;    // time launches tseen % tevery == 0
    int tkind = 0;                // steps per timed launch (1 or 2; 0 before the first): one kind is timed

    int n = 0;  // steps done
};

namespace {

int slot_of(const fdtd_ctx *c, int level) { return ((c->n + 2 - level) % 3 + 3) % 3; }

char *field_dev(fdtd_ctx *c, int field, int level)
{
    const int q = field % 3;
    if (field >= FDTD_HX && field <= FDTD_HZ) return c->H[c->hcur][q];
    const int s = slot_of(c, level);
    return field < 3 ? c->D[s][q] : c->E[s][q];
}

// stored value -> SI value of a field
double field_factor(const fdtd_ctx *c, int field)
{
    return fdtd_field_factor(c->storage, c->mu0, c->eps0, c->sE, c->sD, c->sH, field);
}

int upload_ids(fdtd_ctx *c)
{
    if (c->ids_host.empty()) return 0;
    std::vector<unsigned char> merged(c->ids_host);
    const int nx = c->nx, ny = c->ny;
    for (size_t s = 0; s < c->src.size() / 3; s++) {
        const size_t o = ((size_t)c->src[3 * s + 2] * ny + c->src[3 * s + 1]) * nx + c->src[3 * s];
        merged[o] |= kSourceBit;
    }
    std::vector<unsigned char> img(c->Np);
    fdtd_pitch_bytes(merged.data(), c->nx, (size_t)c->ny * c->nz, c->pitch, img.data());
    // steps queued on any stream (a non-blocking one included) may still read the ids
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(c->ids, img.data(), c->Np, hipMemcpyHostToDevice));
    return 0;
}

// 16-bit scales: explicit ones from FdtdExt, else derived from the source
// waveform (fdtd_auto_scales); fixed from the first step on.
void resolve_scales(fdtd_ctx *c)
{
    if (c->scales_set) return;
    fdtd_auto_scales(c->jz.data(), (int)c->jz.size(), c->eps0, c->ext, &c->sE, &c->sD, &c->sH);
    c->scales_set = true;
    c->lut_dirty = true;
}

int upload_lut(fdtd_ctx *c)
{
    if (!c->lut_dirty) return 0;
    std::vector<MatCoef> lut(c->nmat);
    fdtd_build_lut(c->table.data(), c->nmat, c->dt, c->eps0, c->ext.debye_mode, c->storage, c->sE, c->sD, lut.data());
    // only after set_materials or the first step: queued steps may still read the old table
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(c->lut, lut.data(), sizeof(MatCoef) * c->nmat, hipMemcpyHostToDevice));
    c->lut_dirty = false;
    return 0;
}

void set_shape(fdtd_ctx *c, int ty, int kchunk, int cpl, int txw = 64)
{
    c->ty = ty;
    c->kchunk = std::min(kchunk, c->nz);
    c->cpl = cpl;
    c->txw = cpl == 2 ? 64 : txw;
    const int width = c->cpl == 2 ? 128 : c->txw;
    c->tiles_x = (c->nx + width - 1) / width;
    c->tiles_y = (c->ny + ty - 1) / ty;
    c->tiles_z = (c->nz + c->kchunk - 1) / c->kchunk;
    c->kchunk2 = c->kchunk;
}

// Launch shape a context takes for a grid, from launch-shape sweeps on
// MI355X: 4 tile rows (8 rows gain 1-5% at 256^3 and lose at 320x320x880 and
// on small grids), identity block order; two cells per lane for 16-bit
// fields and the z-chunk where those pay (auto_lane_cells_alone,
// auto_kchunk_alone).  FdtdExt's tile_y, kchunk, lane_cells and tile_x fix
// their parts.  cus = 0: the current device's compute units (256 when none
// can be queried).
void default_shape(int nx, int ny, int nz, const FdtdExt &e, int cus, int *ty, int *kchunk, int *cpl, int *txw)
{
    if (cus <= 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    const size_t esz = fdtd_storage_bytes(e.storage);
    *ty = e.tile_y ? e.tile_y : 4;
    *cpl = auto_lane_cells_alone(nx, ny, nz, *ty, e.kchunk, esz, e.lane_cells, cus);
    *kchunk = e.kchunk ? std::min(e.kchunk, nz) : auto_kchunk_alone(nx, ny, nz, *ty, *cpl, cus);
    *txw = *cpl == 2 ? 64 : (e.tile_x ? e.tile_x : 64);
}

hipError_t upload_murx(fdtd_ctx *c, const float *rpx0, const float *rpxEnd)
{
    const size_t n = (size_t)c->ny * c->nz;
    std::vector<float4> h(2 * n);
    mur_x_coefs(c->ny, c->nz, c->dt, c->dx, c->mu0, c->eps0, rpx0, rpxEnd, h.data(), h.data() + n);
    return hipMemcpy(c->murx, h.data(), sizeof(float4) * 2 * n, hipMemcpyHostToDevice);
}

}  // namespace

extern "C" int fdtd_ext_defaults(FdtdExt *ext)
{
    if (!ext) return FDTD_EINVAL;
    ext->debye_mode = FDTD_DEBYE_FIXED;
    ext->jz_index_mode = FDTD_JZ_STEP;
    ext->jz_limit = 0;
    ext->tile_y = 0;
    ext->kchunk = 0;
    // rows of tiles along x on the 8 XCDs: measured on MI355X against the identity order, 0.7 to
    // 8.4% faster for every grid and storage (x-halo columns stay in one L2); whole layers per XCD
    // (2) and eighths of a layer (4) are slower than the identity at 320x320x880
    ext->xcd_remap = 3;
    ext->storage = FDTD_STORE_F32;
    ext->scale_e = ext->scale_d = ext->scale_h = 0.0f;
    ext->lane_cells = 0;
    ext->tile_x = 0;
    ext->load_zero_terms = 0;
    ext->time_steps = 0;
    return 0;
}

extern "C" int fdtd_default_launch_shape(const FdtdParams *p, const FdtdExt *ext, int cus, int *tile_x, int *tile_y,
                                         int *kchunk, int *lane_cells)
{
    if (!p) return fdtd_fail(FDTD_EINVAL, "null argument");
    if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fdtd_fail(FDTD_ERANGE, "grid %dx%dx%d: each edge must be >= 3", p->nx, p->ny, p->nz);
    FdtdExt e;
    fdtd_ext_defaults(&e);
    if (ext) e = *ext;
    if (int rc = fdtd_check_ext(e)) return rc;
    int ty = 0, kc = 0, cpl = 0, txw = 0;
    default_shape(p->nx, p->ny, p->nz, e, cus, &ty, &kc, &cpl, &txw);
    if (tile_x) *tile_x = cpl == 2 ? 128 : txw;
    if (tile_y) *tile_y = ty;
    if (kchunk) *kchunk = kc;
    if (lane_cells) *lane_cells = cpl;
    return 0;
}

extern "C" int fdtd_ctx_create(const FdtdParams *p, const FdtdExt *ext, fdtd_ctx **out)
{
    if (!p || !out) return fdtd_fail(FDTD_EINVAL, "null argument");
    *out = nullptr;
    if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fdtd_fail(FDTD_ERANGE, "grid %dx%dx%d: each edge must be >= 3", p->nx, p->ny, p->nz);
    const long long N = (long long)p->nx * p->ny * p->nz;
    FdtdExt e;
    fdtd_ext_defaults(&e);
    if (ext) e = *ext;
    if (int rc = fdtd_check_ext(e)) return rc;
    // The fused kernel addresses each field array with 32-bit byte offsets
    // that span the array and its guards.
    const int pitch = fused_pitch(p->nx, fdtd_storage_bytes(e.storage));
    const long long plane = (long long)pitch * p->ny;
    if (plane * p->nz + 2 * fused_guard(pitch, plane) >= (1LL << 30))
        return fdtd_fail(FDTD_ERANGE, "grid of %lld cells: too large for one context", N);

    fdtd_ctx *c = new fdtd_ctx();
    c->nx = p->nx; c->ny = p->ny; c->nz = p->nz;
    c->N = (size_t)N;
    c->pitch = pitch;
    c->plane = plane;
    c->Np = (size_t)plane * p->nz;
    c->dt = p->dt; c->dx = p->dx; c->dy = p->dy; c->dz = p->dz; c->mu0 = p->mu0; c->eps0 = p->eps0;
    c->ext = e;
    c->storage = e.storage;
    c->esz = fdtd_storage_bytes(e.storage);
    int txw = 64;
    default_shape(c->nx, c->ny, c->nz, e, 0, &c->ty, &c->kchunk, &c->cpl, &txw);
    set_shape(c, c->ty, c->kchunk, c->cpl, txw);

    const size_t nf = 24;  // 9 E + 9 D + 6 H
    c->guard = fused_guard(c->pitch, c->plane);
    const size_t blob_elems = nf * (c->Np + kArraySkew) + 2 * (size_t)c->guard;
    hipError_t he = hipMalloc(&c->blob, blob_elems * c->esz);
    if (he != hipSuccess) { delete c; return fdtd_hip_fail(he, "hipMalloc(fields)"); }
    char *q = c->blob + c->guard * c->esz;
    const size_t stride = (c->Np + kArraySkew) * c->esz;
    for (int s = 0; s < 3; s++)
        for (int k = 0; k < 3; k++) { c->E[s][k] = q; q += stride; }
    for (int s = 0; s < 3; s++)
        for (int k = 0; k < 3; k++) { c->D[s][k] = q; q += stride; }
    for (int s = 0; s < 2; s++)
        for (int k = 0; k < 3; k++) { c->H[s][k] = q; q += stride; }
    he = hipMemset(c->blob, 0, blob_elems * c->esz);
    if (he == hipSuccess) he = hipMalloc(&c->ids_base, c->Np + 2 * (size_t)c->guard);
    if (he == hipSuccess) he = hipMemset(c->ids_base, 0, c->Np + 2 * (size_t)c->guard);
    if (he == hipSuccess) c->ids = c->ids_base + c->guard;
    if (he == hipSuccess) he = hipMalloc(&c->lut, sizeof(MatCoef) * kMaxMaterials);
    const size_t rps[6] = {(size_t)2 * c->ny * c->nz, (size_t)c->nx * 2 * c->nz, (size_t)c->nx * c->ny * 2,
                           (size_t)2 * c->ny * c->nz, (size_t)c->nx * 2 * c->nz, (size_t)c->nx * c->ny * 2};
    for (int r = 0; r < 6 && he == hipSuccess; r++) {
        he = hipMalloc(&c->rp[r], rps[r] * sizeof(float));
        if (he == hipSuccess) {
            std::vector<float> ones(rps[r], 1.0f);
            he = hipMemcpy(c->rp[r], ones.data(), rps[r] * sizeof(float), hipMemcpyHostToDevice);
        }
    }
    if (he == hipSuccess) he = hipMalloc(&c->murx, sizeof(float4) * 2 * (size_t)c->ny * c->nz);
    if (he == hipSuccess) {
        std::vector<float> o0((size_t)2 * c->ny * c->nz, 1.0f);
        he = upload_murx(c, o0.data(), o0.data());
    }
    if (he != hipSuccess) {
        int rc = fdtd_hip_fail(he, "fdtd_ctx_create");
        fdtd_ctx_destroy(c);
        return rc;
    }
    // Vacuum default material so a context steps before materials are set.
    c->table = {0.0f, 1.0f, 1.0f, 0.0f};
    c->nmat = 1;
    c->lut_dirty = true;
    *out = c;
    return 0;
}

extern "C" int fdtd_ctx_destroy(fdtd_ctx *c)
{
    if (!c) return 0;
    if (c->side) (void)hipStreamSynchronize(c->side);
    (void)hipFree(c->blob);
    (void)hipFree(c->spare_blob);
    (void)hipFree(c->ids_base);
    (void)hipFree(c->lut);
    for (int r = 0; r < 6; r++) (void)hipFree(c->rp[r]);
    (void)hipFree(c->murx);
    (void)hipFree(c->probes);
    (void)hipFree(c->probe_out);
    (void)hipFree(c->slice_stage);
    (void)hipFree(c->energy_out);
    if (c->slice_host) (void)hipHostFree(c->slice_host);
    if (c->slice_ev) (void)hipEventDestroy(c->slice_ev);
    if (c->side) (void)hipStreamDestroy(c->side);
    for (auto ev : c->tev)
        if (ev) (void)hipEventDestroy(ev);
    delete c;
    return 0;
}

extern "C" int fdtd_ctx_set_materials(fdtd_ctx *c, const unsigned char *ids, int nmat, const float *table)
{
    if (!c || !ids || !table) return fdtd_fail(FDTD_EINVAL, "null argument");
    if (nmat < 1 || nmat > kMaxMaterials) return fdtd_fail(FDTD_ERANGE, "nmat %d not in [1,%d]", nmat, kMaxMaterials);
    for (size_t t = 0; t < c->N; t++)
        if (ids[t] >= nmat) return fdtd_fail(FDTD_ERANGE, "material id %d at cell %zu >= nmat %d", ids[t], t, nmat);
    // the device ids first: the table changes only once they match it
    std::vector<unsigned char> prev(ids, ids + c->N);
    c->ids_host.swap(prev);
    if (int rc = upload_ids(c)) {
        c->ids_host.swap(prev);
        return rc;
    }
    c->table.assign(table, table + 4 * nmat);
    c->nmat = nmat;
    c->lut_dirty = true;
    return 0;
}

extern "C" int fdtd_ctx_set_material_arrays(fdtd_ctx *c, const float *sigma, const float *epsS, const float *epsI,
                                            const float *tauD)
{
    if (!c || !sigma || !epsS || !epsI || !tauD) return fdtd_fail(FDTD_EINVAL, "null argument");
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t>, int> seen;
    std::vector<float> table;
    std::vector<unsigned char> ids(c->N);
    for (size_t t = 0; t < c->N; t++) {
        uint32_t b[4];
        memcpy(&b[0], &sigma[t], 4); memcpy(&b[1], &epsS[t], 4); memcpy(&b[2], &epsI[t], 4); memcpy(&b[3], &tauD[t], 4);
        auto key = std::make_tuple(b[0], b[1], b[2], b[3]);
        auto it = seen.find(key);
        int id;
        if (it == seen.end()) {
            id = (int)seen.size();
            if (id >= kMaxMaterials) return fdtd_fail(FDTD_ERANGE, "more than %d distinct materials", kMaxMaterials);
            seen.emplace(key, id);
            table.insert(table.end(), {sigma[t], epsS[t], epsI[t], tauD[t]});
        } else {
            id = it->second;
        }
        ids[t] = (unsigned char)id;
    }
    return fdtd_ctx_set_materials(c, ids.data(), (int)seen.size(), table.data());
}

extern "C" int fdtd_ctx_set_rp(fdtd_ctx *c, const float *rpx0, const float *rpy0, const float *rpz0,
                               const float *rpxEnd, const float *rpyEnd, const float *rpzEnd)
{
    if (!c) return fdtd_fail(FDTD_EINVAL, "null context");
    const float *src[6] = {rpx0, rpy0, rpz0, rpxEnd, rpyEnd, rpzEnd};
    const size_t rps[6] = {(size_t)2 * c->ny * c->nz, (size_t)c->nx * 2 * c->nz, (size_t)c->nx * c->ny * 2,
                           (size_t)2 * c->ny * c->nz, (size_t)c->nx * 2 * c->nz, (size_t)c->nx * c->ny * 2};
    for (int r = 0; r < 6; r++)
        if (!src[r]) return fdtd_fail(FDTD_EINVAL, "null rp plane %d", r);
    // queued steps (on any stream) may still read the planes and the x-face coefficients
    HIP_TRY(hipDeviceSynchronize());
    for (int r = 0; r < 6; r++)
        HIP_TRY(hipMemcpy(c->rp[r], src[r], rps[r] * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(upload_murx(c, rpx0, rpxEnd));
    return 0;
}

extern "C" int fdtd_ctx_set_sources(fdtd_ctx *c, int nsrc, const int *xyz, int jz_len, const float *jz)
{
    if (!c || nsrc < 0 || jz_len < 0 || (nsrc > 0 && !xyz) || (jz_len > 0 && !jz)) return fdtd_fail(FDTD_EINVAL, "bad sources");
    if (int rc = fdtd_check_sources(c->nx, c->ny, c->nz, nsrc, xyz)) return rc;
    if (c->n == 0) c->scales_set = false;  // auto scales follow the waveform until the first step
    if (c->ids_host.empty()) c->ids_host.assign(c->N, 0);
    c->src.assign(xyz, xyz + 3 * nsrc);
    c->jz.assign(jz, jz + jz_len);
    return upload_ids(c);
}

extern "C" int fdtd_ctx_set_probes(fdtd_ctx *c, int nprobe, const int *probes, int capacity)
{
    if (!c || nprobe < 0 || capacity < 0 || (nprobe > 0 && !probes)) return fdtd_fail(FDTD_EINVAL, "bad probes");
    if (int rc = fdtd_check_probes(c->nx, c->ny, c->nz, nprobe, probes)) return rc;
    HIP_TRY(hipDeviceSynchronize());   // queued Mur launches may still record into the old buffers
    (void)hipFree(c->probes);
    (void)hipFree(c->probe_out);
    c->probes = nullptr;
    c->probe_out = nullptr;
    c->nprobe = 0;                     // stays 0 (no probe launches) unless every step below succeeds
    c->probe_cap = 0;
    c->probe_rows = 0;
    if (nprobe == 0 || capacity == 0) return 0;
    HIP_TRY(hipMalloc(&c->probes, sizeof(int) * 4 * nprobe));
    HIP_TRY(hipMemcpy(c->probes, probes, sizeof(int) * 4 * nprobe, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->probe_out, sizeof(double) * (size_t)nprobe * capacity));
    HIP_TRY(hipMemset(c->probe_out, 0, sizeof(double) * (size_t)nprobe * capacity));
    c->probe_field.resize(nprobe);
    for (int p = 0; p < nprobe; p++) c->probe_field[p] = probes[4 * p];
    c->nprobe = nprobe;
    c->probe_cap = capacity;
    return 0;
}

extern "C" int fdtd_ctx_set_slices(fdtd_ctx *c, int k, int interval, int capacity)
{
    if (!c || interval < 0 || capacity < 0) return fdtd_fail(FDTD_EINVAL, "bad slices");
    if (interval > 0 && (k < 0 || k >= c->nz)) return fdtd_fail(FDTD_ERANGE, "slice plane %d out of range", k);
    if (c->side) HIP_TRY(hipStreamSynchronize(c->side));
    (void)hipFree(c->slice_stage);
    if (c->slice_host) (void)hipHostFree(c->slice_host);
    c->slice_stage = nullptr;
    c->slice_host = nullptr;
    c->slice_k = k;
    c->slice_interval = 0;             // stays 0 (no capture) unless every step below succeeds
    c->slice_cap = 0;
    c->slice_count = 0;
    c->slice_steps.clear();
    if (interval == 0 || capacity == 0) return 0;
    const size_t bytes = (size_t)capacity * 9 * c->plane * c->esz;
    HIP_TRY(hipMalloc(&c->slice_stage, bytes));
    HIP_TRY(hipHostMalloc(&c->slice_host, bytes, hipHostMallocDefault));
    if (!c->side) HIP_TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    if (!c->slice_ev) HIP_TRY(hipEventCreateWithFlags(&c->slice_ev, hipEventDisableTiming));
    c->slice_interval = interval;
    c->slice_cap = capacity;
    return 0;
}

extern "C" int fdtd_ctx_set_field(fdtd_ctx *c, int field, int level, const float *host)
{
    if (!c || !host || field < 0 || field > 8 || level < 0 || level > 2) return fdtd_fail(FDTD_EINVAL, "bad field/level");
    resolve_scales(c);
    std::vector<char> buf(c->Np * c->esz);
    fdtd_to_storage_pitched(c->storage, host, c->nx, (size_t)c->ny * c->nz, c->pitch, field_factor(c, field), buf.data());
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(field_dev(c, field, level), buf.data(), buf.size(), hipMemcpyHostToDevice));
    c->fields_set = true;
    return 0;
}

extern "C" int fdtd_ctx_get_field(fdtd_ctx *c, int field, int level, float *host)
{
    if (!c || !host || field < 0 || field > 8 || level < 0 || level > 2) return fdtd_fail(FDTD_EINVAL, "bad field/level");
    resolve_scales(c);
    std::vector<char> buf(c->Np * c->esz);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(buf.data(), field_dev(c, field, level), buf.size(), hipMemcpyDeviceToHost));
    fdtd_from_storage_pitched(c->storage, buf.data(), c->nx, (size_t)c->ny * c->nz, c->pitch, field_factor(c, field), host);
    return 0;
}

extern "C" int fdtd_ctx_field_ptr(fdtd_ctx *c, int field, int level, void **dev)
{
    if (!c || !dev || field < 0 || field > 8 || level < 0 || level > 2) return fdtd_fail(FDTD_EINVAL, "bad field/level");
    *dev = field_dev(c, field, level);
    return 0;
}

namespace {

// Kernel arguments of step n -> n+1 with the context's one-step launch shape:
// ring slot n % 3 written, (n + 2) % 3 holds level n, (n + 1) % 3 level n-1.
void fill_step_args(fdtd_ctx *c, int n, StepArgs &a)
{
    const int t = n % 3, s1 = (n + 2) % 3, s2 = (n + 1) % 3;
    const int ho = c->hcur, hn = c->hcur ^ 1;
    a.hxo = c->H[ho][0]; a.hyo = c->H[ho][1]; a.hzo = c->H[ho][2];
    a.hxn = c->H[hn][0]; a.hyn = c->H[hn][1]; a.hzn = c->H[hn][2];
    a.ex0 = c->E[s1][0]; a.ey0 = c->E[s1][1]; a.ez0 = c->E[s1][2];
    a.ex1 = c->E[s2][0]; a.ey1 = c->E[s2][1]; a.ez1 = c->E[s2][2];
    a.exT = c->E[t][0]; a.eyT = c->E[t][1]; a.ezT = c->E[t][2];
    a.dx0 = c->D[s1][0]; a.dy0 = c->D[s1][1]; a.dz0 = c->D[s1][2];
    a.dx1 = c->D[s2][0]; a.dy1 = c->D[s2][1]; a.dz1 = c->D[s2][2];
    a.dxT = c->D[t][0]; a.dyT = c->D[t][1]; a.dzT = c->D[t][2];
    a.ids = c->ids;
    a.lut = c->lut;
    a.nmat = c->nmat;
    const float jz = fdtd_jz_at(c->jz.data(), (int)c->jz.size(), c->ext.jz_limit, c->ext.jz_index_mode, n);
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz;
    a.pitch = c->pitch;
    a.plane = c->plane;
    a.jz = fdtd_jz_stored(jz, c->storage, c->eps0, c->sD);
    {
        const GridGeom g{c->dt, c->dx, c->dy, c->dz, c->mu0, c->eps0};
        float ch[3], cd[3];
        fdtd_curl_coefs(g, c->storage, c->sE, c->sD, c->sH, ch, cd);
        a.chx = ch[0]; a.chy = ch[1]; a.chz = ch[2];
        a.cdx = cd[0]; a.cdy = cd[1]; a.cdz = cd[2];
    }
    a.kchunk = c->kchunk;
    a.tiles_x = c->tiles_x; a.tiles_y = c->tiles_y; a.tiles_z = c->tiles_z;
    a.xcd_remap = c->ext.xcd_remap;
    a.guard = (unsigned)c->guard;
    a.mur_x = fused_mur_x(c->nx, c->kchunk, 64 * c->cpl);
    a.cpl = c->cpl;
    a.txw = c->txw;
    a.murx0 = c->murx;
    a.murx1 = c->murx + (size_t)c->ny * c->nz;
    // zero-term rows rely on E^{n-1}, D^{n-1} being finite wherever E^n, D^n are, which stepping from
    // zero (or from any state the steps made) keeps; a state set from outside may not
    a.zskip = (c->storage == FDTD_STORE_F32 && c->ext.load_zero_terms == 0 && !c->fields_set) ? 1 : 0;
}

MurArgs mur_args(const fdtd_ctx *c, void *const *target, void *const *source)
{
    MurArgs m;
    m.exT = target[0]; m.eyT = target[1]; m.ezT = target[2];
    m.exS = source[0]; m.eyS = source[1]; m.ezS = source[2];
    m.rpx0 = c->rp[0]; m.rpy0 = c->rp[1]; m.rpz0 = c->rp[2];
    m.rpxEnd = c->rp[3]; m.rpyEnd = c->rp[4]; m.rpzEnd = c->rp[5];
    m.nx = c->nx; m.ny = c->ny; m.nz = c->nz; m.pitch = c->pitch; m.plane = c->plane;
    m.dt = c->dt; m.hx = c->dx; m.hy = c->dy; m.hz = c->dz; m.mu0 = c->mu0; m.eps0 = c->eps0;
    return m;
}

// The probe row of the step just done, reading the fields given (newest
// levels); p.nprobe = 0 when nothing is recorded.
ProbeArgs probe_row(fdtd_ctx *c, char *const *fields)
{
    ProbeArgs p;
    p.nprobe = 0;
    if (c->nprobe > 0 && c->probe_rows < c->probe_cap) {
        for (int f = 0; f < 9; f++) {
            p.field[f] = fields[f];
            p.factor[f] = field_factor(c, f);
        }
        p.probes = c->probes;
        p.nprobe = c->nprobe;
        p.pitch = c->pitch;
        p.plane = c->plane;
        p.out = c->probe_out;
        p.row = c->probe_rows++;
    }
    return p;
}

// Energy sum and z-slice capture of step c->n, when due.
int after_step(fdtd_ctx *c, hipStream_t stream)
{
    if (c->energy_interval > 0 && c->n % c->energy_interval == 0 && c->energy_rows < c->energy_cap) {
        EnergyArgs g;
        for (int q = 0; q < 3; q++) {
            g.D[q] = field_dev(c, q, 0);
            g.H[q] = field_dev(c, 3 + q, 0);
            g.E[q] = field_dev(c, 6 + q, 0);
        }
        g.nx = c->nx; g.ny = c->ny; g.nz = c->nz; g.pitch = c->pitch; g.plane = c->plane;
        g.ed_factor = field_factor(c, FDTD_EX) * field_factor(c, FDTD_DX);
        g.hh_factor = (double)c->mu0 * field_factor(c, FDTD_HX) * field_factor(c, FDTD_HX);
        g.partial = c->energy_out + c->energy_cap;
        g.out = c->energy_out + c->energy_rows;
        const hipError_t e = launch_energy(g, c->storage, stream);
        if (e != hipSuccess) return fdtd_hip_fail(e, "energy");
        c->energy_steps.push_back(c->n);
        c->energy_rows++;
    }
    if (c->slice_interval > 0 && c->n % c->slice_interval == 0 && c->slice_count < c->slice_cap) {
        const size_t pb = (size_t)c->plane * c->esz;
        char *stage = c->slice_stage + (size_t)c->slice_count * 9 * pb;
        for (int f = 0; f < 9; f++)
            HIP_TRY(hipMemcpyAsync(stage + f * pb, field_dev(c, f, 0) + (size_t)c->slice_k * pb, pb,
                                   hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipEventRecord(c->slice_ev, stream));
        HIP_TRY(hipStreamWaitEvent(c->side, c->slice_ev, 0));
        HIP_TRY(hipMemcpyAsync(c->slice_host + (size_t)c->slice_count * 9 * pb, stage, 9 * pb,
                               hipMemcpyDeviceToHost, c->side));
        c->slice_steps.push_back(c->n);
        c->slice_count++;
    }
    return 0;
}

// Whether to time this launch of `kind` steps: every tevery-th launch of the
// kind timed first, up to the capacity.
bool timed_launch(fdtd_ctx *c, int kind)
{
    if (c->tused >= c->tcap || (c->tkind != 0 && c->tkind != kind)) return false;
    if (c->tseen++ % c->tevery != 0) return false;
    c->tkind = kind;
    return true;
}

// Whether steps n+1 and n+2 may run as one two-step launch: fp32 storage
// stepped from zero (the spare slot's never-written cells then hold what
// the ring slot it replaces holds: zero), no halo exchange between the two,
// and nothing that needs H^{n+1/2} or all fields of step n+1 (H probes, an
// energy sum or a slice at n+1).
bool two_step_ok(const fdtd_ctx *c)
{
    if (c->ext.time_steps != 2 || c->storage != FDTD_STORE_F32 || c->fields_set || c->halo_used) return false;
    if (c->nprobe > 0)
        for (int f : c->probe_field)
            if (f >= FDTD_HX && f <= FDTD_HZ) return false;
    const int m = c->n + 1;
    if (c->energy_interval > 0 && m % c->energy_interval == 0 && c->energy_rows < c->energy_cap) return false;
    if (c->slice_interval > 0 && m % c->slice_interval == 0 && c->slice_count < c->slice_cap) return false;
    return true;
}

int alloc_spare(fdtd_ctx *c)
{
    if (c->spare_blob) return 0;
    const size_t stride = (c->Np + kArraySkew) * c->esz;
    const size_t bytes = 6 * stride + 2 * (size_t)c->guard * c->esz;
    HIP_TRY(hipMalloc(&c->spare_blob, bytes));
    HIP_TRY(hipMemset(c->spare_blob, 0, bytes));
    char *q = c->spare_blob + c->guard * c->esz;
    for (int k = 0; k < 3; k++) { c->Es[k] = q; q += stride; }
    for (int k = 0; k < 3; k++) { c->Ds[k] = q; q += stride; }
    return 0;
}

// Steps n -> n+1 -> n+2 in one fused launch (fused2_tile), then the probe
// row of step n+1 (E and D of level n+1 as stored), step n+2's Mur faces
// and its probe row.
int step_two(fdtd_ctx *c, hipStream_t stream)
{
    resolve_scales(c);
    if (int rc = upload_lut(c)) return rc;
    if (int rc = alloc_spare(c)) return rc;
    const int n = c->n;
    Step2Args s;
    fill_step_args(c, n, s.A);
    const int rows = kTwoStepRows;
    s.A.kchunk = c->kchunk2;
    two_step_tiles(c->nx, c->ny, c->nz, c->kchunk2, rows, &s.A.tiles_x, &s.A.tiles_y, &s.A.tiles_z);
    s.exS = c->Es[0]; s.eyS = c->Es[1]; s.ezS = c->Es[2];
    s.dxS = c->Ds[0]; s.dyS = c->Ds[1]; s.dzS = c->Ds[2];
    const float jz2 = fdtd_jz_at(c->jz.data(), (int)c->jz.size(), c->ext.jz_limit, c->ext.jz_index_mode, n + 1);
    s.jz2 = fdtd_jz_stored(jz2, c->storage, c->eps0, c->sD);
    s.rpy0 = c->rp[1]; s.rpz0 = c->rp[2]; s.rpyEnd = c->rp[4]; s.rpzEnd = c->rp[5];
    s.dt = c->dt; s.hy = c->dy; s.hz = c->dz; s.mu0 = c->mu0; s.eps0 = c->eps0;
    const bool timed = timed_launch(c, 2);
    hipError_t e = timed ? launch_fused_step2(s, rows, stream, c->tev[2 * c->tused], c->tev[2 * c->tused + 1])
                         : launch_fused_step2(s, rows, stream);
    if (e != hipSuccess) return fdtd_hip_fail(e, "fused_step2");
    if (timed) c->tused++;

    // step n+1: level n+1 is in ring slot t; no H probes (two_step_ok)
    const int t = n % 3, tb = (n + 1) % 3;
    c->n = n + 1;
    {
        char *fields[9];
        for (int q = 0; q < 3; q++) {
            fields[q] = c->D[t][q];
            fields[3 + q] = c->H[c->hcur][q];
            fields[6 + q] = c->E[t][q];
        }
        ProbeArgs p = probe_row(c, fields);
        if (p.nprobe > 0) {
            e = launch_probes(p, c->storage, stream);
            if (e != hipSuccess) return fdtd_hip_fail(e, "probes");
        }
    }
    // step n+2: the spare takes ring slot (n+1) % 3 (level n-1 before)
    for (int q = 0; q < 3; q++) {
        std::swap(c->E[tb][q], c->Es[q]);
        std::swap(c->D[tb][q], c->Ds[q]);
    }
    c->n = n + 2;
    c->hcur ^= 1;
    void *tgt[3] = {c->E[tb][0], c->E[tb][1], c->E[tb][2]};
    void *src[3] = {c->E[t][0], c->E[t][1], c->E[t][2]};
    MurArgs m = mur_args(c, tgt, src);
    mur_segments(c->nx, c->ny, c->nz, m.seg_start, 0);
    char *fields[9];
    for (int f = 0; f < 9; f++) fields[f] = field_dev(c, f, 0);
    ProbeArgs p = probe_row(c, fields);
    e = launch_mur_probes(m, p, c->storage, stream);
    if (e != hipSuccess) return fdtd_hip_fail(e, "mur_faces/probes");
    return after_step(c, stream);
}

}  // namespace

extern "C" int fdtd_step_fused(fdtd_ctx *c, hipStream_t stream)
{
    if (!c) return fdtd_fail(FDTD_EINVAL, "null context");
    resolve_scales(c);
    int rc = upload_lut(c);
    if (rc) return rc;
    const int n = c->n;
    StepArgs a;
    fill_step_args(c, n, a);
    const bool timed = timed_launch(c, 1);
    hipError_t e = timed ? launch_fused_step(a, c->ty, c->storage, stream, c->tev[2 * c->tused], c->tev[2 * c->tused + 1])
                         : launch_fused_step(a, c->ty, c->storage, stream);
    if (e != hipSuccess) return fdtd_hip_fail(e, "fused_step");
    if (timed) c->tused++;

    void *tgt[3] = {a.exT, a.eyT, a.ezT};
    void *src[3] = {(void *)a.ex0, (void *)a.ey0, (void *)a.ez0};
    MurArgs m = mur_args(c, tgt, src);
    mur_segments(c->nx, c->ny, c->nz, m.seg_start, mur_skip_mask(a.mur_x));

    c->n = n + 1;
    c->hcur ^= 1;

    // The step's probe row is recorded by the Mur launch (newest levels).
    char *fields[9];
    for (int f = 0; f < 9; f++) fields[f] = field_dev(c, f, 0);
    ProbeArgs p = probe_row(c, fields);
    e = launch_mur_probes(m, p, c->storage, stream);
    if (e != hipSuccess) return fdtd_hip_fail(e, "mur_faces/probes");
    return after_step(c, stream);
}

extern "C" int fdtd_run(fdtd_ctx *c, int steps, hipStream_t stream)
{
    if (!c || steps < 0) return fdtd_fail(FDTD_EINVAL, "bad run");
    for (int s = 0; s < steps;) {
        if (steps - s >= 2 && two_step_ok(c)) {
            if (int rc = step_two(c, stream)) return rc;
            s += 2;
        } else {
            if (int rc = fdtd_step_fused(c, stream)) return rc;
            s += 1;
        }
    }
    return 0;
}

extern "C" int fdtd_ctx_steps_done(fdtd_ctx *c) { return c ? c->n : FDTD_EINVAL; }

extern "C" int fdtd_probe_read(fdtd_ctx *c, double *host, int max_rows)
{
    if (!c || (!host && max_rows > 0)) return fdtd_fail(FDTD_EINVAL, "bad probe read");
    HIP_TRY(hipDeviceSynchronize());
    const int rows = std::min(max_rows, c->probe_rows);
    if (rows > 0 && c->nprobe > 0)
        HIP_TRY(hipMemcpy(host, c->probe_out, sizeof(double) * (size_t)rows * c->nprobe, hipMemcpyDeviceToHost));
    return c->nprobe > 0 ? rows : 0;
}

extern "C" int fdtd_probe_peak_units(fdtd_ctx *c, double *peak, int nprobe)
{
    if (!c || nprobe < 0 || (nprobe > 0 && !peak)) return fdtd_fail(FDTD_EINVAL, "bad probe peak query");
    if (nprobe != c->nprobe) return fdtd_fail(FDTD_EINVAL, "%d probes asked, the context records %d", nprobe, c->nprobe);
    if (nprobe == 0) return 0;
    HIP_TRY(hipDeviceSynchronize());
    const int rows = c->probe_rows;
    if (rows == 0) {   // nothing recorded (and the scales may not be fixed before the first step)
        std::fill(peak, peak + nprobe, 0.0);
        return 0;
    }
    std::vector<double> series((size_t)rows * nprobe);
    HIP_TRY(hipMemcpy(series.data(), c->probe_out, sizeof(double) * series.size(), hipMemcpyDeviceToHost));
    double factor[9];
    for (int f = 0; f < 9; f++) factor[f] = field_factor(c, f);
    return fdtd_probe_peaks(c->storage, series.data(), rows, nprobe, c->probe_field.data(), factor, peak);
}

extern "C" int fdtd_slices_read(fdtd_ctx *c, int *steps, float *data, int max_slices)
{
    if (!c) return fdtd_fail(FDTD_EINVAL, "null context");
    if (c->side) HIP_TRY(hipStreamSynchronize(c->side));
    const int cnt = std::min(max_slices, c->slice_count);
    const size_t pl = (size_t)c->plane, nxy = (size_t)c->nx * c->ny;
    for (int s = 0; s < cnt; s++) {
        if (steps) steps[s] = c->slice_steps[s];
        if (data)
            for (int f = 0; f < 9; f++)
                for (int j = 0; j < c->ny; j++)
                    fdtd_from_storage(c->storage, c->slice_host + (((size_t)s * 9 + f) * pl + (size_t)j * c->pitch) * c->esz,
                                      c->nx, field_factor(c, f), data + ((size_t)s * 9 + f) * nxy + (size_t)j * c->nx);
    }
    return cnt;
}

extern "C" int fdtd_ctx_snapshot(fdtd_ctx *c, float *dev_out, hipStream_t stream)
{
    if (!c || !dev_out) return fdtd_fail(FDTD_EINVAL, "null argument");
    resolve_scales(c);
    SnapArgs a;
    for (int f = 0; f < 9; f++) {
        a.field[f] = field_dev(c, f, 0);
        a.factor[f] = field_factor(c, f);
    }
    a.out = dev_out;
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.pitch = c->pitch; a.plane = c->plane;
    const hipError_t e = launch_snapshot(a, c->storage, stream);
    return e == hipSuccess ? 0 : fdtd_hip_fail(e, "snapshot");
}

extern "C" int fdtd_ctx_launch_shape(fdtd_ctx *c, int *tile_x, int *tile_y, int *kchunk, int *blocks)
{
    if (!c) return fdtd_fail(FDTD_EINVAL, "null context");
    if (tile_x) *tile_x = c->cpl == 2 ? 128 : c->txw;
    if (tile_y) *tile_y = c->ty;
    if (kchunk) *kchunk = c->kchunk;
    if (blocks) *blocks = c->tiles_x * c->tiles_y * c->tiles_z;
    return 0;
}

extern "C" int fdtd_ctx_set_timing(fdtd_ctx *c, int capacity, int every)
{
    if (!c || capacity < 0 || every < 1) return fdtd_fail(FDTD_EINVAL, "bad timing capacity or stride");
    HIP_TRY(hipDeviceSynchronize());
    for (auto ev : c->tev)
        if (ev) (void)hipEventDestroy(ev);
    c->tcap = 0;                       // stays 0 (no timed launches) unless every event is created
    c->tused = 0;
    c->tevery = every;
    c->tseen = 0;
    c->tkind = 0;
    c->tev.assign(2 * (size_t)capacity, nullptr);
    for (auto &ev : c->tev) HIP_TRY(hipEventCreate(&ev));
    c->tcap = capacity;
    return 0;
}

extern "C" int fdtd_ctx_kernel_time(fdtd_ctx *c, double *ms_total, int *launches)
{
    if (!c || !ms_total) return fdtd_fail(FDTD_EINVAL, "null argument");
    double acc = 0.0;
    for (int t = 0; t < c->tused; t++) {
        HIP_TRY(hipEventSynchronize(c->tev[2 * t + 1]));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c->tev[2 * t], c->tev[2 * t + 1]));
        acc += ms;
    }
    *ms_total = acc;
    if (launches) *launches = c->tused;
    return 0;
}

extern "C" int fdtd_ctx_timed_steps(fdtd_ctx *c) { return c ? c->tkind : FDTD_EINVAL; }

extern "C" int fdtd_ctx_scales(fdtd_ctx *c, double *scale_e, double *scale_d, double *scale_h)
{
    if (!c) return fdtd_fail(FDTD_EINVAL, "null context");
    resolve_scales(c);
    if (scale_e) *scale_e = c->sE;
    if (scale_d) *scale_d = c->sD;
    if (scale_h) *scale_h = c->sH;
    return 0;
}

extern "C" int fdtd_ctx_layout(fdtd_ctx *c, int *pitch, long long *plane, int *elem_bytes)
{
    if (!c) return fdtd_fail(FDTD_EINVAL, "null context");
    if (pitch) *pitch = c->pitch;
    if (plane) *plane = c->plane;
    if (elem_bytes) *elem_bytes = (int)c->esz;
    return 0;
}

namespace {

// The planes of one halo message: FDTD_HALO_UP carries E^n (x, y, z) and
// H^{n-1/2} (x, y) of the lower slab's last real plane into the upper
// slab's ghost plane 0; FDTD_HALO_DOWN carries E^n (x, y) of the upper
// slab's first real plane into the lower slab's ghost plane nz-1.
int halo_planes(const fdtd_ctx *c, int dir, bool sending, PlaneCopy &pc, char **fields, long long *plane_index)
{
    static const int up[] = {FDTD_EX, FDTD_EY, FDTD_EZ, FDTD_HX, FDTD_HY};
    static const int down[] = {FDTD_EX, FDTD_EY};
    if (dir != FDTD_HALO_UP && dir != FDTD_HALO_DOWN) return fdtd_fail(FDTD_EINVAL, "halo direction %d", dir);
    if (c->nz < 3) return fdtd_fail(FDTD_ERANGE, "a slab needs a ghost plane and a real plane");
    const int *f = dir == FDTD_HALO_UP ? up : down;
    pc.n = dir == FDTD_HALO_UP ? 5 : 2;
    pc.words = c->plane * (long long)c->esz / 16;
    *plane_index = dir == FDTD_HALO_UP ? (sending ? c->nz - 2 : 0) : (sending ? 1 : c->nz - 1);
    for (int q = 0; q < pc.n; q++) fields[q] = field_dev(const_cast<fdtd_ctx *>(c), f[q], 0);
    return 0;
}

}  // namespace

extern "C" int fdtd_ctx_halo_bytes(fdtd_ctx *c, int dir, size_t *bytes)
{
    if (!c || !bytes) return fdtd_fail(FDTD_EINVAL, "null argument");
    PlaneCopy pc{};
    char *fields[kHaloFields];
    long long k;
    if (int rc = halo_planes(c, dir, true, pc, fields, &k)) return rc;
    *bytes = (size_t)pc.n * (size_t)pc.words * 16;
    return 0;
}

extern "C" int fdtd_ctx_halo_pack(fdtd_ctx *c, int dir, void *dst, hipStream_t stream)
{
    if (!c || !dst) return fdtd_fail(FDTD_EINVAL, "null argument");
    if ((uintptr_t)dst % 16) return fdtd_fail(FDTD_EINVAL, "halo buffer must be 16-byte aligned");
    PlaneCopy pc{};
    char *fields[kHaloFields];
    long long k;
    if (int rc = halo_planes(c, dir, true, pc, fields, &k)) return rc;
    c->halo_used = true;
    const size_t pb = (size_t)pc.words * 16;
    for (int q = 0; q < pc.n; q++) {
        pc.src[q] = (const uint4 *)(fields[q] + (size_t)k * pb);
        pc.dst[q] = (uint4 *)((char *)dst + q * pb);
    }
    const hipError_t e = launch_plane_copy(pc, stream);
    return e == hipSuccess ? 0 : fdtd_hip_fail(e, "halo pack");
}

extern "C" int fdtd_ctx_halo_unpack(fdtd_ctx *c, int dir, const void *src, hipStream_t stream)
{
    if (!c || !src) return fdtd_fail(FDTD_EINVAL, "null argument");
    if ((uintptr_t)src % 16) return fdtd_fail(FDTD_EINVAL, "halo buffer must be 16-byte aligned");
    PlaneCopy pc{};
    char *fields[kHaloFields];
    long long k;
    if (int rc = halo_planes(c, dir, false, pc, fields, &k)) return rc;
    c->halo_used = true;
    const size_t pb = (size_t)pc.words * 16;
    for (int q = 0; q < pc.n; q++) {
        pc.src[q] = (const uint4 *)((const char *)src + q * pb);
        pc.dst[q] = (uint4 *)(fields[q] + (size_t)k * pb);
    }
    // Zero-term rows stay on: the message holds cells a neighbour's fused
    // kernel stepped (fdtd_ctx_halo_pack), so a non-finite E^{n-1} or D^{n-1}
    // in the ghost plane still implies a non-finite E^n or D^n there, and the
    // ghost plane's own results are overwritten by the next exchange.
    const hipError_t e = launch_plane_copy(pc, stream);
    return e == hipSuccess ? 0 : fdtd_hip_fail(e, "halo unpack");
}

extern "C" int fdtd_ctx_set_energy(fdtd_ctx *c, int interval, int capacity)
{
    if (!c || interval < 0 || capacity < 0) return fdtd_fail(FDTD_EINVAL, "bad energy series");
    HIP_TRY(hipDeviceSynchronize());
    (void)hipFree(c->energy_out);
    c->energy_out = nullptr;
    c->energy_interval = 0;            // stays 0 (no sums) unless the allocation succeeds
    c->energy_cap = 0;
    c->energy_rows = 0;
    c->energy_steps.clear();
    if (interval == 0 || capacity == 0) return 0;
    HIP_TRY(hipMalloc(&c->energy_out, sizeof(double) * ((size_t)capacity + kEnergyBlocks)));
    c->energy_interval = interval;
    c->energy_cap = capacity;
    return 0;
}

extern "C" int fdtd_energy_read(fdtd_ctx *c, int *steps, double *energy, int max_rows)
{
    if (!c || max_rows < 0) return fdtd_fail(FDTD_EINVAL, "bad energy read");
    HIP_TRY(hipDeviceSynchronize());
    const int rows = std::min(max_rows, c->energy_rows);
    if (rows > 0 && energy) HIP_TRY(hipMemcpy(energy, c->energy_out, sizeof(double) * rows, hipMemcpyDeviceToHost));
    for (int r = 0; r < rows && steps; r++) steps[r] = c->energy_steps[r];
    return rows;
}

// Binary snapshot: a fixed header, then the 24 field arrays in the device
// layout and storage type (E slots, D slots, H buffers; pitched rows),
// written as they are so a restore is bit-exact.  Materials, sources, rp
// planes, probes and slices are inputs, not state: a restoring caller sets
// them as before.
namespace {

struct SnapHeader {
    char magic[8];            // "FDTDSNP1"
    int nx, ny, nz, pitch;
    long long plane;
    int storage, esz, n, hcur;
    int debye_mode, jz_index_mode;
    double sE, sD, sH;
    int scales_set, pad;
};

}  // namespace

extern "C" int fdtd_ctx_save(fdtd_ctx *c, const char *path)
{
    if (!c || !path) return fdtd_fail(FDTD_EINVAL, "null argument");
    resolve_scales(c);
    HIP_TRY(hipDeviceSynchronize());
    SnapHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, "FDTDSNP1", 8);
    h.nx = c->nx; h.ny = c->ny; h.nz = c->nz; h.pitch = c->pitch; h.plane = c->plane;
    h.storage = c->storage; h.esz = (int)c->esz; h.n = c->n; h.hcur = c->hcur;
    h.debye_mode = c->ext.debye_mode; h.jz_index_mode = c->ext.jz_index_mode;
    h.sE = c->sE; h.sD = c->sD; h.sH = c->sH; h.scales_set = c->scales_set ? 1 : 0;
    FILE *f = fopen(path, "wb");
    if (!f) return fdtd_fail(FDTD_EIO, "cannot write %s", path);
    bool ok = fwrite(&h, sizeof h, 1, f) == 1;
    std::vector<char> buf(c->Np * c->esz);
    char *const *arrays[3] = {&c->E[0][0], &c->D[0][0], &c->H[0][0]};
    const int counts[3] = {9, 9, 6};
    for (int g = 0; g < 3 && ok; g++)
        for (int q = 0; q < counts[g] && ok; q++) {
            if (hipMemcpy(buf.data(), arrays[g][q], buf.size(), hipMemcpyDeviceToHost) != hipSuccess) ok = false;
            else ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
        }
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : fdtd_fail(FDTD_EIO, "writing snapshot %s failed", path);
}

extern "C" int fdtd_ctx_load(fdtd_ctx *c, const char *path)
{
    if (!c || !path) return fdtd_fail(FDTD_EINVAL, "null argument");
    FILE *f = fopen(path, "rb");
    if (!f) return fdtd_fail(FDTD_EIO, "cannot read %s", path);
    SnapHeader h;
    if (fread(&h, sizeof h, 1, f) != 1 || memcmp(h.magic, "FDTDSNP1", 8) != 0) {
        fclose(f);
        return fdtd_fail(FDTD_EIO, "%s is not a snapshot", path);
    }
    if (h.nx != c->nx || h.ny != c->ny || h.nz != c->nz || h.pitch != c->pitch || h.plane != c->plane ||
        h.storage != c->storage || h.esz != (int)c->esz) {
        fclose(f);
        return fdtd_fail(FDTD_ERANGE, "snapshot %s is of another grid or storage", path);
    }
    if (h.n < 0 || (h.hcur != 0 && h.hcur != 1) || !(h.sE > 0.0 && h.sD > 0.0 && h.sH > 0.0) ||
        !std::isfinite(h.sE) || !std::isfinite(h.sD) || !std::isfinite(h.sH)) {
        fclose(f);
        return fdtd_fail(FDTD_EIO, "snapshot %s has a corrupt header", path);
    }
    // the whole file is checked before any device field changes
    const long long want = (long long)sizeof h + 24LL * (long long)(c->Np * c->esz);
    if (fseeko(f, 0, SEEK_END) != 0 || (long long)ftello(f) != want || fseeko(f, (off_t)sizeof h, SEEK_SET) != 0) {
        fclose(f);
        return fdtd_fail(FDTD_EIO, "snapshot %s is not %lld bytes long", path, want);
    }
    HIP_TRY(hipDeviceSynchronize());
    std::vector<char> buf(c->Np * c->esz);
    char *const *arrays[3] = {&c->E[0][0], &c->D[0][0], &c->H[0][0]};
    const int counts[3] = {9, 9, 6};
    bool ok = true;
    for (int g = 0; g < 3 && ok; g++)
        for (int q = 0; q < counts[g] && ok; q++) {
            ok = fread(buf.data(), 1, buf.size(), f) == buf.size();
            if (ok && hipMemcpy(arrays[g][q], buf.data(), buf.size(), hipMemcpyHostToDevice) != hipSuccess) ok = false;
        }
    fclose(f);
    if (!ok) return fdtd_fail(FDTD_EIO, "reading snapshot %s failed", path);
    c->n = h.n;
    c->hcur = h.hcur;
    c->sE = h.sE; c->sD = h.sD; c->sH = h.sH;
    c->scales_set = h.scales_set != 0;
    c->lut_dirty = true;
    c->fields_set = true;
    return 0;
}

// Times each launch shape on the context's own grid for trial_steps steps
// (after two untimed ones), keeps the fastest and returns the context to
// its zero state at step 0.
extern "C" int fdtd_ctx_autotune(fdtd_ctx *c, int trial_steps, hipStream_t stream, int *tile_y, int *kchunk,
                                 int *lane_cells)
{
    if (!c || trial_steps < 1) return fdtd_fail(FDTD_EINVAL, "bad autotune");
    if (c->n != 0 || c->fields_set) return fdtd_fail(FDTD_ESTATE, "autotune needs an unstepped context with zero fields");
    HIP_TRY(hipDeviceSynchronize());
    const int slice_interval = c->slice_interval, energy_interval = c->energy_interval, tcap = c->tcap;
    c->slice_interval = 0;
    c->energy_interval = 0;
    c->tcap = 0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipError_t e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    int rc = e == hipSuccess ? 0 : fdtd_hip_fail(e, "hipEventCreate");
    // z-chunks 3-7 land within about 1% of each other on fp32 at 256^3 and
    // 320x320x880 (3 best at the first, 4-6 at the second): the trials pick
    const int tys[3] = {2, 4, 8}, kcs[5] = {3, 4, 6, 8, 16};
    // (cells per lane, threads per row): single cells 64 or 128 wide, 16-bit pairs 128 cells wide
    const int forms[3][2] = {{1, 64}, {1, 128}, {2, 64}};
    const int nform = c->esz == 2 ? 3 : 2;
    float best = 0.0f;
    int bty = c->ty, bkc = c->kchunk, bcpl = c->cpl, btxw = c->txw;
    for (int a = 0; a < 3 && !rc; a++)
        for (int b = 0; b < 5 && !rc; b++)
            for (int f = 0; f < nform && !rc; f++) {
                const int q = forms[f][0], w = forms[f][1];
                set_shape(c, tys[a], kcs[b], q, w);
                rc = fdtd_run(c, 2, stream);
                if (rc) break;
                e = hipEventRecord(t0, stream);
                if (e == hipSuccess) rc = fdtd_run(c, trial_steps, stream);
                if (!rc && e == hipSuccess) e = hipEventRecord(t1, stream);
                if (!rc && e == hipSuccess) e = hipEventSynchronize(t1);
                float ms = 0.0f;
                if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
                if (!rc && e != hipSuccess) rc = fdtd_hip_fail(e, "autotune timing");
                if (!rc && (best == 0.0f || ms < best)) {
                    best = ms;
                    bty = tys[a];
                    bkc = kcs[b];
                    bcpl = q;
                    btxw = w;
                }
            }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    // back to the zero state at step 0
    e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipMemset(c->blob, 0, (24 * (c->Np + kArraySkew) + 2 * (size_t)c->guard) * c->esz);
    if (e == hipSuccess && c->spare_blob)
        e = hipMemset(c->spare_blob, 0, (6 * (c->Np + kArraySkew) + 2 * (size_t)c->guard) * c->esz);
    if (!rc && e != hipSuccess) rc = fdtd_hip_fail(e, "autotune reset");
    c->n = 0;
    c->hcur = 0;
    c->probe_rows = 0;
    c->slice_interval = slice_interval;
    c->energy_interval = energy_interval;
    c->tcap = tcap;
    c->tused = 0;
    c->tseen = 0;
    c->tkind = 0;
    set_shape(c, bty, bkc, bcpl, btxw);
    if (rc) return rc;
    if (tile_y) *tile_y = bty;
    if (kchunk) *kchunk = bkc;
    if (lane_cells) *lane_cells = bcpl;
    return 0;
}
