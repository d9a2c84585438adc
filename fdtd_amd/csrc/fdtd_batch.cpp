// Batches of independent grids (fdtd_amd.h, fdtd_batch_*): one device arena
// holds every grid's state; each batch step is one persistent fused launch
// per storage form present, one Mur launch and one probe launch over all
// running grids.  The launches read the grids' descriptors, the per-form
// tile lists and a step counter from device memory, so a sequence of steps
// is captured once per (tile rows, forms present, length) as a HIP graph and
// replayed while grids join and leave; the host re-uploads the tables
// (stream-ordered, from pinned staging) only when the running set changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "fdtd_batch_internal.h"

namespace {

constexpr size_t kAlign = 4096;

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Device tables read by the batch kernels.
struct DevTables {
    BatchInst inst[kBatchMaxInst];
    BatchClass cls[kBatchClasses];
    BatchShared sh;
};

// Byte layout of one grid's arena region.
struct Layout {
    size_t N, Np, esz, skew;       // cells, elements per field array, bytes per element, array stagger
    int pitch;
    long long plane;
    long long guard;
    size_t blob, blob_bytes, ids, lut, jz, rp[6], murx, probes, out, out_bytes, total;
    int jz_n;
};

Layout layout_of(int nx, int ny, int nz, int storage, int nmat, int jz_n, int nprobe, int steps)
{
    Layout L;
    L.N = (size_t)nx * ny * nz;
    L.esz = fdtd_storage_bytes(storage);
    L.pitch = fused_pitch(nx, L.esz);
    L.plane = (long long)L.pitch * ny;
    L.Np = (size_t)L.plane * nz;
    // Channel stagger between the 24 field arrays of large grids (as in a
    // single context); small grids are not worth the memory.
    L.skew = L.Np >= (size_t)1 << 20 ? 65536 : 0;
    L.guard = fused_guard(L.pitch, L.plane);
    L.jz_n = jz_n;
    size_t o = 0;
    L.blob = o;
    L.blob_bytes = (24 * (L.Np + L.skew) + 2 * (size_t)L.guard) * L.esz;
    o = round_up(o + L.blob_bytes, 256);
    L.ids = o;
    o = round_up(o + L.Np + 2 * (size_t)L.guard, 256);
    L.lut = o;
    o = round_up(o + (size_t)nmat * sizeof(MatCoef), 256);
    L.jz = o;
    o = round_up(o + (size_t)jz_n * sizeof(float), 256);
    const size_t rps[6] = {(size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2,
                           (size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2};
    for (int r = 0; r < 6; r++) {
        L.rp[r] = o;
        o = round_up(o + rps[r] * sizeof(float), 256);
    }
    L.murx = o;
    o = round_up(o + 2 * (size_t)ny * nz * sizeof(float4), 256);
    L.probes = o;
    o = round_up(o + (size_t)nprobe * 4 * sizeof(int), 256);
    L.out = o;
    L.out_bytes = (size_t)nprobe * steps * sizeof(double);
    o = round_up(o + L.out_bytes, 256);
    L.total = round_up(o, kAlign);
    return L;
}

int jz_entries(const FdtdExt &e, int jz_len)
{
    return e.jz_limit > 0 ? std::min(e.jz_limit, jz_len) : jz_len;
}

int check_instance(const FdtdParams *p, const FdtdExt &e, const FdtdInstance *in)
{
    if (!p || !in) return fdtd_fail(FDTD_EINVAL, "null argument");
    if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fdtd_fail(FDTD_ERANGE, "grid %dx%dx%d: each edge must be >= 3", p->nx, p->ny, p->nz);
    if (int rc = fdtd_check_ext(e)) return rc;
    const int pitch = fused_pitch(p->nx, fdtd_storage_bytes(e.storage));
    const long long plane = (long long)pitch * p->ny;
    if (plane * p->nz + 2 * fused_guard(pitch, plane) >= (1LL << 30))
        return fdtd_fail(FDTD_ERANGE, "grid of %lld cells: too large", (long long)p->nx * p->ny * p->nz);
    if (in->nmat < 1 || in->nmat > kMaxMaterials || !in->table) return fdtd_fail(FDTD_ERANGE, "nmat %d not in [1,%d]", in->nmat, kMaxMaterials);
    if (in->steps < 1) return fdtd_fail(FDTD_ERANGE, "steps %d", in->steps);
    if (in->nsrc < 0 || (in->nsrc > 0 && !in->src_xyz) || in->jz_len < 0 || (in->jz_len > 0 && !in->jz))
        return fdtd_fail(FDTD_EINVAL, "bad sources");
    if (in->nprobe < 0 || in->nprobe > kBatchMaxProbes || (in->nprobe > 0 && !in->probes))
        return fdtd_fail(FDTD_ERANGE, "nprobe %d not in [0,%d]", in->nprobe, kBatchMaxProbes);
    if (int rc = fdtd_check_sources(p->nx, p->ny, p->nz, in->nsrc, in->src_xyz)) return rc;
    if (int rc = fdtd_check_probes(p->nx, p->ny, p->nz, in->nprobe, in->probes)) return rc;
    return 0;
}

FdtdExt ext_or_defaults(const FdtdExt *ext)
{
    FdtdExt e;
    fdtd_ext_defaults(&e);
    if (ext) e = *ext;
    return e;
}

}  // namespace

// First-fit allocator over the batch's one device allocation; free ranges
// are kept sorted and coalesced.
bool Arena::alloc(size_t len, size_t *off)
{
    len = round_up(len, kAlign);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < len) continue;
        *off = it->first;
        const size_t rest = it->second - len, o2 = it->first + len;
        free_.erase(it);
        if (rest) free_[o2] = rest;
        used_ += len;
        return true;
    }
    return false;
}

void Arena::release(size_t off, size_t len)
{
    len = round_up(len, kAlign);
    used_ -= len;
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

size_t Arena::largest() const
{
    size_t m = 0;
    for (auto &f : free_) m = std::max(m, f.second);
    return m;
}

struct fdtd_batch {
    char *arena_base = nullptr;
    Arena arena;
    DevTables *dev = nullptr;
    DevTables *stage[2] = {nullptr, nullptr};  // pinned
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_pending[2] = {false, false};
    int stage_next = 0;
    hipStream_t up = nullptr;   // region uploads
    hipStream_t cap = nullptr;  // graph capture
    int cus = 0;

    struct Inst {
        bool used = false;
        size_t off = 0;
        Layout L;
        BatchInst d;
        long long start = 0;
        int steps = 0;
        double sE = 1, sD = 1, sH = 1;
        float mu0 = 0, eps0 = 0;
        std::vector<int> probe_field;  // field of each probe
    };
    std::vector<Inst> in;
    long long step = 0;
    int ty = 4, kchunk = 8, graph_steps = 8;
    bool dirty = true;                 // device tables out of date
    std::vector<int> running;          // slots in the device tables
    int mask = 0;                      // kernel classes in the device tables
    int grid[kBatchClasses] = {};
    int tiles[kBatchClasses] = {};
    long long cells[kBatchClasses] = {};
    std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;
    hipStream_t last = nullptr;
    bool ran = false;
};

namespace {

BatchLaunch launch_of(fdtd_batch *b, int cls, int off)
{
    BatchLaunch l;
    l.inst = b->dev->inst;
    l.cls = &b->dev->cls[cls];
    l.sh = &b->dev->sh;
    l.off = off;
    l.xcd_runs = kBatchXcdRun;
    return l;
}

int fused_grid(fdtd_batch *b, int cls)
{
    const int occ = batch_fused_occupancy(b->ty, cls);
    return std::max(1, occ) * b->cus;
}

bool is_running(const fdtd_batch *b, int s)
{
    const auto &I = b->in[s];
    return I.used && b->step - I.start < I.steps;
}

// Rebuilds the work tables for the grids running at the current step and
// enqueues their upload on `stream` when they differ from the device's.
int sync_tables(fdtd_batch *b, hipStream_t stream)
{
    std::vector<int> run;
    for (int s = 0; s < (int)b->in.size(); s++)
        if (is_running(b, s)) run.push_back(s);
    if (!b->dirty && run == b->running) return 0;

    const int k = b->stage_next;
    if (b->stage_pending[k]) {
        HIP_TRY(hipEventSynchronize(b->stage_ev[k]));
        b->stage_pending[k] = false;
    }
    DevTables *T = b->stage[k];
    int maxslot = -1;
    for (int c = 0; c < kBatchClasses; c++) {
        T->cls[c].n = 0;
        T->cls[c].total = 0;
        T->cls[c].start[0] = 0;
        b->tiles[c] = 0;
        b->cells[c] = 0;
    }
    BatchShared &S = T->sh;
    S.step = (int)b->step;
    S.nmur = 0;
    S.mur_total = 0;
    S.mur_start[0] = 0;
    S.nrec = 0;
    int mask = 0;
    for (int s : run) {
        auto &I = b->in[s];
        BatchInst &d = I.d;
        d.kchunk = std::min(b->kchunk, d.nz);
        const int tw = 64 * d.cpl;
        d.mur_x = fused_mur_x(d.nx, d.kchunk, tw);
        mur_segments(d.nx, d.ny, d.nz, d.mur_seg, mur_skip_mask(d.mur_x));
        d.tiles_x = (d.nx + tw - 1) / tw;
        d.tiles_y = (d.ny + b->ty - 1) / b->ty;
        d.tiles_z = (d.nz + d.kchunk - 1) / d.kchunk;
        T->inst[s] = d;
        maxslot = std::max(maxslot, s);
        const int c = batch_class(d.storage, d.cpl);
        mask |= 1 << c;
        BatchClass &C = T->cls[c];
        const int nt = d.tiles_x * d.tiles_y * d.tiles_z;
        C.idx[C.n] = s;
        C.total += nt;
        C.start[++C.n] = C.total;
        b->tiles[c] += nt;
        b->cells[c] += (long long)d.nx * d.ny * d.nz;
        const int mb = (d.mur_seg[12] + kMurBlock - 1) / kMurBlock;
        S.mur_idx[S.nmur] = s;
        S.mur_total += mb;
        S.mur_start[++S.nmur] = S.mur_total;
        for (int p = 0; p < d.nprobe; p++) {
            S.rec[S.nrec][0] = s;
            S.rec[S.nrec][1] = p;
            S.nrec++;
        }
    }
    if (maxslot >= 0)
        HIP_TRY(hipMemcpyAsync(b->dev->inst, T->inst, sizeof(BatchInst) * (maxslot + 1), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(b->dev->cls, T->cls, sizeof(T->cls), hipMemcpyHostToDevice, stream));
    const size_t head = offsetof(BatchShared, rec) + sizeof(S.rec[0]) * (size_t)S.nrec;
    HIP_TRY(hipMemcpyAsync(&b->dev->sh, &S, head, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipEventRecord(b->stage_ev[k], stream));
    b->stage_pending[k] = true;
    b->stage_next = k ^ 1;
    b->running = run;
    b->mask = mask;
    b->dirty = false;
    return 0;
}

int graph_for(fdtd_batch *b, int mask, int g, hipGraphExec_t *out)
{
    const auto key = std::make_tuple(b->ty, mask, g);
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        *out = it->second;
        return 0;
    }
    int grids[kBatchClasses];
    for (int c = 0; c < kBatchClasses; c++) grids[c] = (mask >> c & 1) ? fused_grid(b, c) : 0;
    const int mur_grid = 4 * b->cus;
    HIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int s = 0; s < g && e == hipSuccess; s++) {
        for (int c = 0; c < kBatchClasses && e == hipSuccess; c++)
            if (mask >> c & 1) e = launch_fused_batch(launch_of(b, c, s), b->ty, c, grids[c], b->cap);
        if (e == hipSuccess) e = launch_mur_batch(launch_of(b, 0, s), mur_grid, b->cap);  // and probe rows
    }
    if (e == hipSuccess) e = launch_batch_advance(&b->dev->sh, g, b->cap);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->cap, &graph);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fdtd_hip_fail(e, "batch launch under capture");
    }
    if (e2 != hipSuccess) return fdtd_hip_fail(e2, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fdtd_hip_fail(e, "hipGraphInstantiate");
    for (int c = 0; c < kBatchClasses; c++)
        if (grids[c]) b->grid[c] = grids[c];
    b->graphs[key] = exec;
    *out = exec;
    return 0;
}

void drop_graphs(fdtd_batch *b)
{
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
    b->graphs.clear();
}

int wait_idle(fdtd_batch *b)
{
    if (b->ran) HIP_TRY(hipDeviceSynchronize());
    b->ran = false;
    return 0;
}

}  // namespace

extern "C" int fdtd_batch_create(size_t hbm_bytes, fdtd_batch **out)
{
    if (!out) return fdtd_fail(FDTD_EINVAL, "null argument");
    *out = nullptr;
    if (hbm_bytes < kAlign) return fdtd_fail(FDTD_ERANGE, "arena of %zu bytes", hbm_bytes);
    fdtd_batch *b = new fdtd_batch();
    hbm_bytes = round_up(hbm_bytes, kAlign);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipMalloc(&b->arena_base, hbm_bytes);
    if (e == hipSuccess) e = hipMalloc(&b->dev, sizeof(DevTables));
    if (e == hipSuccess) e = hipMemset(b->dev, 0, sizeof(DevTables));
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
        e = hipHostMalloc(&b->stage[k], sizeof(DevTables), hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->stage_ev[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->up, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    if (e != hipSuccess) {
        const int rc = fdtd_hip_fail(e, "fdtd_batch_create");
        fdtd_batch_destroy(b);
        return rc;
    }
    b->arena.init(hbm_bytes);
    b->in.resize(kBatchMaxInst);
    *out = b;
    return 0;
}

extern "C" int fdtd_batch_destroy(fdtd_batch *b)
{
    if (!b) return 0;
    (void)hipDeviceSynchronize();
    drop_graphs(b);
    (void)hipFree(b->arena_base);
    (void)hipFree(b->dev);
    for (int k = 0; k < 2; k++) {
        if (b->stage[k]) (void)hipHostFree(b->stage[k]);
        if (b->stage_ev[k]) (void)hipEventDestroy(b->stage_ev[k]);
    }
    if (b->up) (void)hipStreamDestroy(b->up);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    delete b;
    return 0;
}

extern "C" int fdtd_batch_instance_bytes(const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *inst,
                                         size_t *bytes)
{
    if (!p || !inst || !bytes) return fdtd_fail(FDTD_EINVAL, "null argument");
    const FdtdExt e = ext_or_defaults(ext);
    if (int rc = check_instance(p, e, inst)) return rc;
    const Layout L = layout_of(p->nx, p->ny, p->nz, e.storage, inst->nmat, jz_entries(e, inst->jz_len), inst->nprobe,
                               inst->steps);
    *bytes = L.total;
    return 0;
}

size_t batch_largest_free(fdtd_batch *b) { return b->arena.largest(); }
size_t batch_used_bytes(fdtd_batch *b) { return b->arena.used(); }
long long batch_step(fdtd_batch *b) { return b->step; }

extern "C" int fdtd_batch_add(fdtd_batch *b, const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *in, int *id)
{
    if (!b || !id) return fdtd_fail(FDTD_EINVAL, "null argument");
    const FdtdExt e = ext_or_defaults(ext);
    if (int rc = check_instance(p, e, in)) return rc;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const size_t N = (size_t)nx * ny * nz;
    if (in->ids)
        for (size_t t = 0; t < N; t++)
            if (in->ids[t] >= in->nmat) return fdtd_fail(FDTD_ERANGE, "material id %d at cell %zu >= nmat %d", in->ids[t], t, in->nmat);
    if (b->step + in->steps >= INT_MAX) return fdtd_fail(FDTD_ERANGE, "batch step counter would overflow");
    int slot = -1;
    for (int s = 0; s < (int)b->in.size(); s++)
        if (!b->in[s].used) { slot = s; break; }
    if (slot < 0) return fdtd_fail(FDTD_ENOMEM, "all %d batch slots in use", kBatchMaxInst);
    const int jz_n = jz_entries(e, in->jz_len);
    const Layout L = layout_of(nx, ny, nz, e.storage, in->nmat, jz_n, in->nprobe, in->steps);
    size_t off = 0;
    if (!b->arena.alloc(L.total, &off))
        return fdtd_fail(FDTD_ENOMEM, "grid needs %zu bytes; largest free region %zu", L.total, b->arena.largest());

    auto &I = b->in[slot];
    I.off = off;
    I.L = L;
    I.steps = in->steps;
    I.mu0 = p->mu0;
    I.eps0 = p->eps0;
    I.probe_field.resize(in->nprobe);
    for (int q = 0; q < in->nprobe; q++) I.probe_field[q] = in->probes[4 * q];
    if (e.storage != kStoreF32) fdtd_auto_scales(in->jz, in->jz_len, p->eps0, e, &I.sE, &I.sD, &I.sH);
    else I.sE = I.sD = I.sH = 1.0;

    char *base = b->arena_base + off;
    // ids with source bits, coefficient table, stored waveform, Mur planes, probes
    std::vector<unsigned char> ids(L.Np, 0);
    if (in->ids) fdtd_pitch_bytes(in->ids, nx, (size_t)ny * nz, L.pitch, ids.data());
    for (int s = 0; s < in->nsrc; s++) {
        const int *q = in->src_xyz + 3 * s;
        ids[(size_t)q[2] * L.plane + (size_t)q[1] * L.pitch + q[0]] |= kSourceBit;
    }
    std::vector<MatCoef> lut(in->nmat);
    fdtd_build_lut(in->table, in->nmat, p->dt, p->eps0, e.debye_mode, e.storage, I.sE, I.sD, lut.data());
    std::vector<float> jz(jz_n);
    for (int t = 0; t < jz_n; t++) jz[t] = fdtd_jz_stored(in->jz[t], e.storage, p->eps0, I.sD);
    const size_t rps[6] = {(size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2,
                           (size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2};
    std::vector<float> ones;
    // x-face Mur coefficients from the rp planes (or vacuum)
    std::vector<float4> murx(2 * (size_t)ny * nz);
    {
        std::vector<float> one_x((size_t)2 * ny * nz, 1.0f);
        mur_x_coefs(ny, nz, p->dt, p->dx, p->mu0, p->eps0, in->rp[0] ? in->rp[0] : one_x.data(),
                    in->rp[3] ? in->rp[3] : one_x.data(), murx.data(), murx.data() + (size_t)ny * nz);
    }
    hipError_t he = hipMemsetAsync(base, 0, L.ids + L.Np + 2 * (size_t)L.guard, b->up);
    if (he == hipSuccess)
        he = hipMemcpyAsync(base + L.murx, murx.data(), sizeof(float4) * murx.size(), hipMemcpyHostToDevice, b->up);
    if (he == hipSuccess) he = hipMemsetAsync(base + L.out, 0, L.out_bytes, b->up);
    if (he == hipSuccess) he = hipMemcpyAsync(base + L.ids + L.guard, ids.data(), L.Np, hipMemcpyHostToDevice, b->up);
    if (he == hipSuccess) he = hipMemcpyAsync(base + L.lut, lut.data(), sizeof(MatCoef) * in->nmat, hipMemcpyHostToDevice, b->up);
    if (he == hipSuccess && jz_n > 0) he = hipMemcpyAsync(base + L.jz, jz.data(), sizeof(float) * jz_n, hipMemcpyHostToDevice, b->up);
    for (int r = 0; r < 6 && he == hipSuccess; r++) {
        const float *src = in->rp[r];
        if (!src) {
            ones.assign(rps[r], 1.0f);
            src = ones.data();
        }
        he = hipMemcpyAsync(base + L.rp[r], src, rps[r] * sizeof(float), hipMemcpyHostToDevice, b->up);
        if (he == hipSuccess && !in->rp[r]) he = hipStreamSynchronize(b->up);  // `ones` is reused
    }
    if (he == hipSuccess && in->nprobe > 0)
        he = hipMemcpyAsync(base + L.probes, in->probes, sizeof(int) * 4 * in->nprobe, hipMemcpyHostToDevice, b->up);
    if (he == hipSuccess) he = hipStreamSynchronize(b->up);
    if (he != hipSuccess) {
        (void)hipStreamSynchronize(b->up);   // copies queued before the failure still read the host vectors
        b->arena.release(off, L.total);
        return fdtd_hip_fail(he, "fdtd_batch_add upload");
    }

    BatchInst &d = I.d;
    memset(&d, 0, sizeof d);
    char *blob = base + L.blob + (size_t)L.guard * L.esz;
    const size_t stride = (L.Np + L.skew) * L.esz;
    for (int s = 0; s < 3; s++)
        for (int q = 0; q < 3; q++) d.E[s][q] = blob + (size_t)(s * 3 + q) * stride;
    for (int s = 0; s < 3; s++)
        for (int q = 0; q < 3; q++) d.D[s][q] = blob + (size_t)(9 + s * 3 + q) * stride;
    for (int s = 0; s < 2; s++)
        for (int q = 0; q < 3; q++) d.H[s][q] = blob + (size_t)(18 + s * 3 + q) * stride;
    d.ids = (const uint8_t *)(base + L.ids + L.guard);
    d.lut = (const MatCoef *)(base + L.lut);
    d.jz = (const float *)(base + L.jz);
    for (int r = 0; r < 6; r++) d.rp[r] = (const float *)(base + L.rp[r]);
    d.probes = (const int *)(base + L.probes);
    d.probe_out = (double *)(base + L.out);
    for (int f = 0; f < 9; f++) d.factor[f] = fdtd_field_factor(e.storage, p->mu0, p->eps0, I.sE, I.sD, I.sH, f);
    d.plane = L.plane;
    d.nmat = in->nmat;
    d.jz_len = jz_n;
    d.jz_hold3 = e.jz_index_mode == FDTD_JZ_HOLD3;
    d.start = (int)b->step;
    d.nsteps = in->steps;
    d.storage = e.storage;
    d.cpl = auto_lane_cells(nx, L.esz, e.lane_cells);
    d.nx = nx; d.ny = ny; d.nz = nz; d.pitch = L.pitch;
    d.guard = (unsigned)L.guard;
    const GridGeom g{p->dt, p->dx, p->dy, p->dz, p->mu0, p->eps0};
    fdtd_curl_coefs(g, e.storage, I.sE, I.sD, I.sH, d.ch, d.cd);
    d.nprobe = in->nprobe;
    d.probe_rows = in->steps;
    d.dt = p->dt; d.hx = p->dx; d.hy = p->dy; d.hz = p->dz; d.mu0 = p->mu0; d.eps0 = p->eps0;
    d.murx0 = (const float4 *)(base + L.murx);
    d.murx1 = d.murx0 + (size_t)ny * nz;
    // zero-term rows: exact fp32 form; batch grids start from zero fields
    d.zskip = (e.storage == FDTD_STORE_F32 && e.load_zero_terms == 0) ? 1 : 0;
    I.start = b->step;
    I.used = true;
    b->dirty = true;
    *id = slot;
    return 0;
}

int batch_remove_nosync(fdtd_batch *b, int id)
{
    if (!b || id < 0 || id >= (int)b->in.size() || !b->in[id].used) return fdtd_fail(FDTD_EINVAL, "bad instance id %d", id);
    auto &I = b->in[id];
    b->arena.release(I.off, I.L.total);
    I.used = false;
    b->dirty = true;
    return 0;
}

extern "C" int fdtd_batch_remove(fdtd_batch *b, int id)
{
    if (!b) return fdtd_fail(FDTD_EINVAL, "null batch");
    if (int rc = wait_idle(b)) return rc;
    return batch_remove_nosync(b, id);
}

extern "C" int fdtd_batch_run(fdtd_batch *b, int steps, hipStream_t stream)
{
    if (!b || steps < 0) return fdtd_fail(FDTD_EINVAL, "bad run");
    if (b->step + steps >= INT_MAX) return fdtd_fail(FDTD_ERANGE, "batch step counter would overflow");
    int left = steps;
    while (left > 0) {
        const int g = std::min(left, b->graph_steps);
        if (int rc = sync_tables(b, stream)) return rc;
        if (b->running.empty()) {
            // nothing runs: the counter moves on the host; the next table
            // upload carries it to the device
            b->step += left;
            b->dirty = true;
            break;
        }
        hipGraphExec_t exec = nullptr;
        if (int rc = graph_for(b, b->mask, g, &exec)) return rc;
        HIP_TRY(hipGraphLaunch(exec, stream));
        b->ran = true;
        b->last = stream;
        b->step += g;
        left -= g;
    }
    return 0;
}

extern "C" int fdtd_batch_status(fdtd_batch *b, int id, int *steps_done, int *steps_total)
{
    if (!b || id < 0 || id >= (int)b->in.size() || !b->in[id].used) return fdtd_fail(FDTD_EINVAL, "bad instance id %d", id);
    const auto &I = b->in[id];
    const long long n = std::min<long long>(std::max<long long>(b->step - I.start, 0), I.steps);
    if (steps_done) *steps_done = (int)n;
    if (steps_total) *steps_total = I.steps;
    return 0;
}

int batch_probe_copy_async(fdtd_batch *b, int id, double *host, int rows, hipStream_t stream)
{
    const auto &I = b->in[id];
    if (rows > 0 && I.d.nprobe > 0)
        HIP_TRY(hipMemcpyAsync(host, I.d.probe_out, sizeof(double) * (size_t)rows * I.d.nprobe, hipMemcpyDeviceToHost,
                               stream));
    return 0;
}

extern "C" int fdtd_batch_probe_read(fdtd_batch *b, int id, double *host, int max_rows)
{
    if (!b || id < 0 || id >= (int)b->in.size() || !b->in[id].used) return fdtd_fail(FDTD_EINVAL, "bad instance id %d", id);
    if (!host && max_rows > 0) return fdtd_fail(FDTD_EINVAL, "null buffer");
    if (int rc = wait_idle(b)) return rc;
    int done = 0;
    fdtd_batch_status(b, id, &done, nullptr);
    const int rows = std::min(max_rows, done);
    const auto &I = b->in[id];
    if (rows > 0 && I.d.nprobe > 0)
        HIP_TRY(hipMemcpy(host, I.d.probe_out, sizeof(double) * (size_t)rows * I.d.nprobe, hipMemcpyDeviceToHost));
    return I.d.nprobe > 0 ? rows : 0;
}

extern "C" int fdtd_batch_probe_peak_units(fdtd_batch *b, int id, double *peak, int nprobe)
{
    if (!b || id < 0 || id >= (int)b->in.size() || !b->in[id].used) return fdtd_fail(FDTD_EINVAL, "bad instance id %d", id);
    const auto &I = b->in[id];
    if (nprobe < 0 || (nprobe > 0 && !peak)) return fdtd_fail(FDTD_EINVAL, "bad probe peak query");
    if (nprobe != I.d.nprobe) return fdtd_fail(FDTD_EINVAL, "%d probes asked, the grid records %d", nprobe, I.d.nprobe);
    if (nprobe == 0) return 0;
    std::vector<double> series((size_t)I.steps * nprobe);
    const int rows = fdtd_batch_probe_read(b, id, series.data(), I.steps);
    if (rows < 0) return rows;
    return fdtd_probe_peaks(I.d.storage, series.data(), rows, nprobe, I.probe_field.data(), I.d.factor, peak);
}

extern "C" int fdtd_batch_get_field(fdtd_batch *b, int id, int field, int level, float *host)
{
    if (!b || id < 0 || id >= (int)b->in.size() || !b->in[id].used) return fdtd_fail(FDTD_EINVAL, "bad instance id %d", id);
    if (!host || field < 0 || field > 8 || level < 0 || level > 2) return fdtd_fail(FDTD_EINVAL, "bad field/level");
    if (int rc = wait_idle(b)) return rc;
    int n = 0;
    fdtd_batch_status(b, id, &n, nullptr);
    const auto &I = b->in[id];
    const void *src;
    if (field >= FDTD_HX && field <= FDTD_HZ) {
        src = I.d.H[n & 1][field - 3];
    } else {
        const int s = ((n + 2 - level) % 3 + 3) % 3;
        src = field < 3 ? I.d.D[s][field] : I.d.E[s][field - 6];
    }
    std::vector<char> buf(I.L.Np * I.L.esz);
    HIP_TRY(hipMemcpy(buf.data(), src, buf.size(), hipMemcpyDeviceToHost));
    fdtd_from_storage_pitched(I.d.storage, buf.data(), I.d.nx, (size_t)I.d.ny * I.d.nz, I.L.pitch, I.d.factor[field], host);
    return 0;
}

extern "C" int fdtd_batch_set_shape(fdtd_batch *b, int tile_y, int kchunk, int graph_steps)
{
    if (!b) return fdtd_fail(FDTD_EINVAL, "null batch");
    if (tile_y != 0 && tile_y != 2 && tile_y != 4 && tile_y != 8) return fdtd_fail(FDTD_EINVAL, "tile_y %d", tile_y);
    if (kchunk < 0 || graph_steps < 0) return fdtd_fail(FDTD_EINVAL, "negative kchunk/graph_steps");
    if (tile_y) b->ty = tile_y;
    if (kchunk) b->kchunk = kchunk;
    if (graph_steps) b->graph_steps = graph_steps;
    b->dirty = true;
    return 0;
}

extern "C" int fdtd_batch_autotune(fdtd_batch *b, int trial_steps, hipStream_t stream, int *tile_y, int *kchunk)
{
    if (!b || trial_steps < 1) return fdtd_fail(FDTD_EINVAL, "bad autotune");
    int count = 0;
    for (auto &I : b->in) {
        if (!I.used) continue;
        if (I.start != b->step) return fdtd_fail(FDTD_ESTATE, "autotune needs unstepped grids");
        count++;
    }
    if (count == 0) return fdtd_fail(FDTD_ESTATE, "autotune needs grids");
    if (int rc = wait_idle(b)) return rc;
    const long long step0 = b->step;
    // Every grid runs through all trials; probe rows stay bounded by the
    // grid's own step count (probe_rows), and its step count is restored.
    std::vector<int> saved(b->in.size(), 0);
    for (size_t s = 0; s < b->in.size(); s++) {
        auto &I = b->in[s];
        if (!I.used) continue;
        saved[s] = I.steps;
        I.steps = I.d.nsteps = INT_MAX / 4;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipError_t e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    int rc = e == hipSuccess ? 0 : fdtd_hip_fail(e, "hipEventCreate");
    const int tys[3] = {2, 4, 8}, kcs[5] = {3, 4, 6, 8, 16};
    float best = 0.0f;
    int bty = b->ty, bkc = b->kchunk;
    for (int a = 0; a < 3 && !rc; a++)
        for (int c = 0; c < 5 && !rc; c++) {
            b->ty = tys[a];
            b->kchunk = kcs[c];
            b->dirty = true;
            rc = fdtd_batch_run(b, 2, stream);  // builds the graphs, uploads the tables
            if (rc) break;
            e = hipEventRecord(t0, stream);
            if (e == hipSuccess) rc = fdtd_batch_run(b, trial_steps, stream);
            if (!rc && e == hipSuccess) e = hipEventRecord(t1, stream);
            if (!rc && e == hipSuccess) e = hipEventSynchronize(t1);
            float ms = 0.0f;
            if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
            if (!rc && e != hipSuccess) rc = fdtd_hip_fail(e, "autotune timing");
            if (!rc && (best == 0.0f || ms < best)) {
                best = ms;
                bty = tys[a];
                bkc = kcs[c];
            }
        }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    // Back to the zero initial state at step0 with the fastest shape.
    e = hipStreamSynchronize(stream);
    if (!rc && e != hipSuccess) rc = fdtd_hip_fail(e, "autotune");
    for (size_t s = 0; s < b->in.size(); s++) {
        auto &I = b->in[s];
        if (!I.used) continue;
        I.steps = I.d.nsteps = saved[s];
        char *base = b->arena_base + I.off;
        e = hipMemsetAsync(base + I.L.blob, 0, I.L.blob_bytes, b->up);
        if (e == hipSuccess) e = hipMemsetAsync(base + I.L.out, 0, I.L.out_bytes, b->up);
        if (!rc && e != hipSuccess) rc = fdtd_hip_fail(e, "autotune reset");
    }
    e = hipStreamSynchronize(b->up);
    if (!rc && e != hipSuccess) rc = fdtd_hip_fail(e, "autotune reset");
    b->step = step0;
    b->ty = bty;
    b->kchunk = bkc;
    b->dirty = true;
    b->ran = false;
    if (rc) return rc;
    if (tile_y) *tile_y = bty;
    if (kchunk) *kchunk = bkc;
    return 0;
}

extern "C" int fdtd_batch_info(fdtd_batch *b, FdtdBatchInfo *info)
{
    if (!b || !info) return fdtd_fail(FDTD_EINVAL, "null argument");
    memset(info, 0, sizeof *info);
    for (int s = 0; s < (int)b->in.size(); s++) {
        if (!b->in[s].used) continue;
        info->instances++;
        if (is_running(b, s)) info->running++;
    }
    info->step = b->step;
    info->arena_bytes = b->arena.size();
    info->used_bytes = b->arena.used();
    info->largest_free = b->arena.largest();
    info->tile_y = b->ty;
    info->kchunk = b->kchunk;
    info->graph_steps = b->graph_steps;
    for (int c = 0; c < kBatchClasses; c++) {  // per storage form
        const int st = class_storage(c);
        info->grid[st] = std::max(info->grid[st], fused_grid(b, c));
        info->tiles[st] += b->tiles[c];
        info->cells[st] += b->cells[c];
    }
    return 0;
}
