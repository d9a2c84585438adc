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
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "fdtd_internal.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail((int)e, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

#define HIP_TRY(call)                                    \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

// Host float <-> 16-bit conversions, round to nearest even.
uint16_t f32_to_f16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);  // NaN
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); // rounds past 65504: inf
    if (ax <= 0x33000000u) return (uint16_t)sign;              // at most half the smallest subnormal
    const int e = (int)(ax >> 23) - 127;
    const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    if (e < -14) {  // subnormal half: value = r * 2^-24
        const int shift = -e - 1;  // 24-bit mantissa (2^e scaled) to units of 2^-24
        const uint32_t half = 1u << (shift - 1);
        uint32_t r = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1);
        if (rem > half || (rem == half && (r & 1u))) r++;
        return (uint16_t)(sign | r);
    }
    uint32_t r = ((uint32_t)(e + 15) << 10) | ((m >> 13) & 0x3ffu);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) r++;
    return (uint16_t)(sign | r);
}

float f16_to_f32(uint16_t h)
{
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float v;
    if (e == 0) {
        v = std::ldexp((float)m, -24);
    } else if (e == 31) {
        const uint32_t x = 0x7f800000u | (m << 13);
        memcpy(&v, &x, 4);
    } else {
        const uint32_t x = ((e + 112) << 23) | (m << 13);
        memcpy(&v, &x, 4);
    }
    return (h & 0x8000u) ? -v : v;
}

uint16_t f32_to_bf16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    return (uint16_t)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}

float bf16_to_f32(uint16_t h)
{
    const uint32_t x = (uint32_t)h << 16;
    float f;
    memcpy(&f, &x, 4);
    return f;
}

}  // namespace

MatCoef fdtd_mat_coef(float dDt, float dEps0, float sg, float es, float ei, float tau, int debye_mode)
{
    // Same expressions and precisions as fdtd_calculations.cu:141-155.
    MatCoef m;
    m.a = 2.0 * dEps0 * ei * tau;
    const float b = m.a + 2.0 * dDt * (dEps0 * es + sg * tau);
    m.c = 1.0 / (b + sg * dDt * dDt);
    m.d = (debye_mode == FDTD_DEBYE_REF) ? (float)(2.0 * b - sg * dDt * dDt)
                                         : (float)((double)m.a + (double)b - (double)(sg * dDt * dDt));
    m.pad = 0.0f;
    m.g1 = 2.0 * (dDt + tau);
    m.g2 = 2.0 * dDt + 4.0 * tau;
    m.g3 = 2.0 * tau;
    m.k1 = m.k2 = m.k3 = m.k5 = 0.0f;
    return m;
}

namespace {

// Scaled-unit E coefficients (see MatCoef), from the same a, b, c, d in
// double; ratio = sD / sE.
void scaled_coef(MatCoef &m, double dt, double eps0, double sg, double es, double ei, double tau, int debye_mode,
                 double ratio)
{
    const double a = 2.0 * eps0 * ei * tau;
    const double b = a + 2.0 * dt * (eps0 * es + sg * tau);
    const double c = 1.0 / (b + sg * dt * dt);
    const double d = (debye_mode == FDTD_DEBYE_REF ? 2.0 * b : a + b) - sg * dt * dt;
    m.k1 = (float)(c * d);
    m.k2 = (float)(c * a);
    m.k3 = (float)(c * eps0 * ratio * 2.0 * dt);
    m.k5 = (float)(c * eps0 * ratio * 2.0 * tau);
}

// Extra elements between consecutive field arrays in the blob: staggers the
// arrays' starts across HBM channels (a no-compute 15-read/9-write stream
// probe measured ~4% more bandwidth with 64 Ki elements of stagger).
constexpr size_t kArraySkew = 65536;

}  // namespace

struct fdtd_ctx {
    int nx, ny, nz;
    size_t N;
    int pitch;
    long long plane;
    float dt, dx, dy, dz, mu0, eps0;
    FdtdExt ext;
    int storage;
    size_t esz;                          // bytes per stored element
    int ty, kchunk, tiles_x, tiles_y, tiles_z;

    long long guard = 0;                 // elements before/after the field arrays and the ids
    char *blob = nullptr;                // all field arrays
    unsigned char *ids_base = nullptr;   // allocation holding ids with its guards
    char *E[3][3], *D[3][3], *H[2][3];   // [slot][component], H[buffer][component]
    int hcur = 0;                        // H buffer holding the newest H
    unsigned char *ids = nullptr;
    MatCoef *lut = nullptr;
    int nmat = 0;
    std::vector<float> table;             // nmat x {sigma, eps_s, eps_inf, tau}
    bool lut_dirty = true;
    float *rp[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    std::vector<unsigned char> ids_host;  // material ids without source bits
    std::vector<int> src;                 // 3 per source
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

    std::vector<hipEvent_t> tev;  // start/end pairs around fused launches
    int tcap = 0, tused = 0;

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
    if (c->storage == kStoreF32) return 1.0;
    const double eta0 = std::sqrt((double)c->mu0 / (double)c->eps0);
    if (field < 3) return (double)c->eps0 * c->sD;
    if (field < 6) return c->sH / eta0;
    return c->sE;
}

void to_storage(const fdtd_ctx *c, const float *in, size_t n, double factor, std::vector<char> &out)
{
    out.resize(n * c->esz);
    if (c->storage == kStoreF32) {
        memcpy(out.data(), in, n * 4);
        return;
    }
    if (c->storage == kStoreF32Fast) {
        float *o = (float *)out.data();
        for (size_t t = 0; t < n; t++) o[t] = (float)((double)in[t] / factor);
        return;
    }
    uint16_t *o = (uint16_t *)out.data();
    for (size_t t = 0; t < n; t++) {
        const float v = (float)((double)in[t] / factor);
        o[t] = c->storage == kStoreF16 ? f32_to_f16(v) : f32_to_bf16(v);
    }
}

void from_storage(const fdtd_ctx *c, const char *in, size_t n, double factor, float *out)
{
    if (c->storage == kStoreF32) {
        memcpy(out, in, n * 4);
        return;
    }
    if (c->storage == kStoreF32Fast) {
        const float *f = (const float *)in;
        for (size_t t = 0; t < n; t++) out[t] = (float)((double)f[t] * factor);
        return;
    }
    const uint16_t *h = (const uint16_t *)in;
    for (size_t t = 0; t < n; t++) {
        const float v = c->storage == kStoreF16 ? f16_to_f32(h[t]) : bf16_to_f32(h[t]);
        out[t] = (float)((double)v * factor);
    }
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
    HIP_TRY(hipMemcpy(c->ids, merged.data(), c->N, hipMemcpyHostToDevice));
    return 0;
}

// 16-bit scales: explicit ones from FdtdExt, else 1/32 of the largest
// per-step D/eps0 kick of the source waveform: fields near a source then sit
// at O(10^2-10^3) in storage units, far below the fp16 overflow (65504) and
// with about 10^7 of range above the fp16 subnormals.  Fixed from the first
// step on.
void resolve_scales(fdtd_ctx *c)
{
    if (c->scales_set) return;
    double peak = 0.0;
    for (float v : c->jz) peak = std::max(peak, (double)std::fabs(v));
    const double s = peak > 0.0 ? peak / (double)c->eps0 / 32.0 : 1.0;
    c->sE = c->ext.scale_e > 0.0f ? c->ext.scale_e : s;
    c->sD = c->ext.scale_d > 0.0f ? c->ext.scale_d : s;
    c->sH = c->ext.scale_h > 0.0f ? c->ext.scale_h : s;
    c->scales_set = true;
    c->lut_dirty = true;
}

int upload_lut(fdtd_ctx *c)
{
    if (!c->lut_dirty) return 0;
    std::vector<MatCoef> lut(c->nmat);
    for (int m = 0; m < c->nmat; m++) {
        const float *t = &c->table[4 * m];
        lut[m] = fdtd_mat_coef(c->dt, c->eps0, t[0], t[1], t[2], t[3], c->ext.debye_mode);
        if (c->storage != kStoreF32)
            scaled_coef(lut[m], c->dt, c->eps0, t[0], t[1], t[2], t[3], c->ext.debye_mode, c->sD / c->sE);
    }
    HIP_TRY(hipMemcpy(c->lut, lut.data(), sizeof(MatCoef) * c->nmat, hipMemcpyHostToDevice));
    c->lut_dirty = false;
    return 0;
}

}  // namespace

extern "C" const char *fdtd_last_error(void) { return g_err.c_str(); }

extern "C" int fdtd_ext_defaults(FdtdExt *ext)
{
    if (!ext) return FDTD_EINVAL;
    ext->debye_mode = FDTD_DEBYE_FIXED;
    ext->jz_index_mode = FDTD_JZ_STEP;
    ext->jz_limit = 0;
    ext->tile_y = 0;
    ext->kchunk = 0;
    ext->xcd_remap = 0;  // fastest measured; 2 cuts fabric re-reads to 1.10x but runs ~5% slower
    ext->storage = FDTD_STORE_F32;
    ext->scale_e = ext->scale_d = ext->scale_h = 0.0f;
    return 0;
}

extern "C" int fdtd_ctx_create(const FdtdParams *p, const FdtdExt *ext, fdtd_ctx **out)
{
    if (!p || !out) return fail(FDTD_EINVAL, "null argument");
    *out = nullptr;
    if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fail(FDTD_ERANGE, "grid %dx%dx%d: each edge must be >= 3", p->nx, p->ny, p->nz);
    const long long N = (long long)p->nx * p->ny * p->nz;
    // The fused kernel addresses each field array with 32-bit byte offsets
    // that span the array and its guards.
    if (N + 2 * fused_guard(p->nx, (long long)p->nx * p->ny) >= (1LL << 30))
        return fail(FDTD_ERANGE, "grid of %lld cells: too large for one context", N);
    FdtdExt e;
    fdtd_ext_defaults(&e);
    if (ext) e = *ext;
    if (e.debye_mode != FDTD_DEBYE_REF && e.debye_mode != FDTD_DEBYE_FIXED) return fail(FDTD_EINVAL, "debye_mode %d", e.debye_mode);
    if (e.jz_index_mode != FDTD_JZ_STEP && e.jz_index_mode != FDTD_JZ_HOLD3) return fail(FDTD_EINVAL, "jz_index_mode %d", e.jz_index_mode);
    if (e.tile_y != 0 && e.tile_y != 2 && e.tile_y != 4 && e.tile_y != 8) return fail(FDTD_EINVAL, "tile_y %d", e.tile_y);
    if (e.kchunk < 0 || e.jz_limit < 0) return fail(FDTD_EINVAL, "negative kchunk/jz_limit");
    if (e.xcd_remap < 0 || e.xcd_remap > 2) return fail(FDTD_EINVAL, "xcd_remap %d", e.xcd_remap);
    if (e.storage != FDTD_STORE_F32 && e.storage != FDTD_STORE_F16 && e.storage != FDTD_STORE_BF16 &&
        e.storage != FDTD_STORE_F32_FAST)
        return fail(FDTD_EINVAL, "storage %d", e.storage);
    if (!(e.scale_e >= 0.0f && e.scale_d >= 0.0f && e.scale_h >= 0.0f)) return fail(FDTD_EINVAL, "negative scale");

    fdtd_ctx *c = new fdtd_ctx();
    c->nx = p->nx; c->ny = p->ny; c->nz = p->nz;
    c->N = (size_t)N;
    c->pitch = p->nx;
    c->plane = (long long)p->nx * p->ny;
    c->dt = p->dt; c->dx = p->dx; c->dy = p->dy; c->dz = p->dz; c->mu0 = p->mu0; c->eps0 = p->eps0;
    c->ext = e;
    c->storage = e.storage;
    c->esz = (e.storage == FDTD_STORE_F32 || e.storage == FDTD_STORE_F32_FAST) ? 4 : 2;
    // Defaults from launch-shape sweeps on MI355X (256^3 and 320x320x880
    // Debye phantoms): 64x4 tiles marching 8 planes, identity block order.
    c->ty = e.tile_y ? e.tile_y : 4;
    c->tiles_x = (c->nx + 63) / 64;
    c->tiles_y = (c->ny + c->ty - 1) / c->ty;
    c->kchunk = e.kchunk ? std::min(e.kchunk, c->nz) : std::min(8, c->nz);
    c->tiles_z = (c->nz + c->kchunk - 1) / c->kchunk;

    const size_t nf = 24;  // 9 E + 9 D + 6 H
    c->guard = fused_guard(c->pitch, c->plane);
    const size_t blob_elems = nf * (c->N + kArraySkew) + 2 * (size_t)c->guard;
    hipError_t he = hipMalloc(&c->blob, blob_elems * c->esz);
    if (he != hipSuccess) { delete c; return hip_fail(he, "hipMalloc(fields)"); }
    char *q = c->blob + c->guard * c->esz;
    const size_t stride = (c->N + kArraySkew) * c->esz;
    for (int s = 0; s < 3; s++)
        for (int k = 0; k < 3; k++) { c->E[s][k] = q; q += stride; }
    for (int s = 0; s < 3; s++)
        for (int k = 0; k < 3; k++) { c->D[s][k] = q; q += stride; }
    for (int s = 0; s < 2; s++)
        for (int k = 0; k < 3; k++) { c->H[s][k] = q; q += stride; }
    he = hipMemset(c->blob, 0, blob_elems * c->esz);
    if (he == hipSuccess) he = hipMalloc(&c->ids_base, c->N + 2 * (size_t)c->guard);
    if (he == hipSuccess) he = hipMemset(c->ids_base, 0, c->N + 2 * (size_t)c->guard);
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
    if (he != hipSuccess) {
        int rc = hip_fail(he, "fdtd_ctx_create");
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
    (void)hipFree(c->ids_base);
    (void)hipFree(c->lut);
    for (int r = 0; r < 6; r++) (void)hipFree(c->rp[r]);
    (void)hipFree(c->probes);
    (void)hipFree(c->probe_out);
    (void)hipFree(c->slice_stage);
    if (c->slice_host) (void)hipHostFree(c->slice_host);
    if (c->slice_ev) (void)hipEventDestroy(c->slice_ev);
    if (c->side) (void)hipStreamDestroy(c->side);
    for (auto ev : c->tev) (void)hipEventDestroy(ev);
    delete c;
    return 0;
}

extern "C" int fdtd_ctx_set_materials(fdtd_ctx *c, const unsigned char *ids, int nmat, const float *table)
{
    if (!c || !ids || !table) return fail(FDTD_EINVAL, "null argument");
    if (nmat < 1 || nmat > kMaxMaterials) return fail(FDTD_ERANGE, "nmat %d not in [1,%d]", nmat, kMaxMaterials);
    for (size_t t = 0; t < c->N; t++)
        if (ids[t] >= nmat) return fail(FDTD_ERANGE, "material id %d at cell %zu >= nmat %d", ids[t], t, nmat);
    c->table.assign(table, table + 4 * nmat);
    c->nmat = nmat;
    c->lut_dirty = true;
    c->ids_host.assign(ids, ids + c->N);
    return upload_ids(c);
}

extern "C" int fdtd_ctx_set_material_arrays(fdtd_ctx *c, const float *sigma, const float *epsS, const float *epsI,
                                            const float *tauD)
{
    if (!c || !sigma || !epsS || !epsI || !tauD) return fail(FDTD_EINVAL, "null argument");
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
            if (id >= kMaxMaterials) return fail(FDTD_ERANGE, "more than %d distinct materials", kMaxMaterials);
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
    if (!c) return fail(FDTD_EINVAL, "null context");
    const float *src[6] = {rpx0, rpy0, rpz0, rpxEnd, rpyEnd, rpzEnd};
    const size_t rps[6] = {(size_t)2 * c->ny * c->nz, (size_t)c->nx * 2 * c->nz, (size_t)c->nx * c->ny * 2,
                           (size_t)2 * c->ny * c->nz, (size_t)c->nx * 2 * c->nz, (size_t)c->nx * c->ny * 2};
    for (int r = 0; r < 6; r++) {
        if (!src[r]) return fail(FDTD_EINVAL, "null rp plane %d", r);
        HIP_TRY(hipMemcpy(c->rp[r], src[r], rps[r] * sizeof(float), hipMemcpyHostToDevice));
    }
    return 0;
}

extern "C" int fdtd_ctx_set_sources(fdtd_ctx *c, int nsrc, const int *xyz, int jz_len, const float *jz)
{
    if (!c || nsrc < 0 || jz_len < 0 || (nsrc > 0 && !xyz) || (jz_len > 0 && !jz)) return fail(FDTD_EINVAL, "bad sources");
    for (int s = 0; s < nsrc; s++) {
        const int x = xyz[3 * s], y = xyz[3 * s + 1], z = xyz[3 * s + 2];
        if (x < 1 || x >= c->nx - 1 || y < 1 || y >= c->ny - 1 || z < 0 || z >= c->nz - 1)
            return fail(FDTD_ERANGE, "source %d at (%d,%d,%d) outside the Dz update range", s, x, y, z);
    }
    if (c->n == 0) c->scales_set = false;  // auto scales follow the waveform until the first step
    if (c->ids_host.empty()) c->ids_host.assign(c->N, 0);
    c->src.assign(xyz, xyz + 3 * nsrc);
    c->jz.assign(jz, jz + jz_len);
    return upload_ids(c);
}

extern "C" int fdtd_ctx_set_probes(fdtd_ctx *c, int nprobe, const int *probes, int capacity)
{
    if (!c || nprobe < 0 || capacity < 0 || (nprobe > 0 && !probes)) return fail(FDTD_EINVAL, "bad probes");
    for (int p = 0; p < nprobe; p++) {
        const int *q = probes + 4 * p;
        if (q[0] < 0 || q[0] > 8 || q[1] < 0 || q[1] >= c->nx || q[2] < 0 || q[2] >= c->ny || q[3] < 0 || q[3] >= c->nz)
            return fail(FDTD_ERANGE, "probe %d out of range", p);
    }
    (void)hipFree(c->probes);
    (void)hipFree(c->probe_out);
    c->probes = nullptr;
    c->probe_out = nullptr;
    c->nprobe = nprobe;
    c->probe_cap = capacity;
    c->probe_rows = 0;
    if (nprobe == 0 || capacity == 0) { c->nprobe = 0; return 0; }
    HIP_TRY(hipMalloc(&c->probes, sizeof(int) * 4 * nprobe));
    HIP_TRY(hipMemcpy(c->probes, probes, sizeof(int) * 4 * nprobe, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->probe_out, sizeof(double) * (size_t)nprobe * capacity));
    HIP_TRY(hipMemset(c->probe_out, 0, sizeof(double) * (size_t)nprobe * capacity));
    return 0;
}

extern "C" int fdtd_ctx_set_slices(fdtd_ctx *c, int k, int interval, int capacity)
{
    if (!c || interval < 0 || capacity < 0) return fail(FDTD_EINVAL, "bad slices");
    if (interval > 0 && (k < 0 || k >= c->nz)) return fail(FDTD_ERANGE, "slice plane %d out of range", k);
    if (c->side) HIP_TRY(hipStreamSynchronize(c->side));
    (void)hipFree(c->slice_stage);
    if (c->slice_host) (void)hipHostFree(c->slice_host);
    c->slice_stage = nullptr;
    c->slice_host = nullptr;
    c->slice_k = k;
    c->slice_interval = interval;
    c->slice_cap = capacity;
    c->slice_count = 0;
    c->slice_steps.clear();
    if (interval == 0 || capacity == 0) { c->slice_interval = 0; return 0; }
    const size_t bytes = (size_t)capacity * 9 * c->plane * c->esz;
    HIP_TRY(hipMalloc(&c->slice_stage, bytes));
    HIP_TRY(hipHostMalloc(&c->slice_host, bytes, hipHostMallocDefault));
    if (!c->side) HIP_TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    if (!c->slice_ev) HIP_TRY(hipEventCreateWithFlags(&c->slice_ev, hipEventDisableTiming));
    return 0;
}

extern "C" int fdtd_ctx_set_field(fdtd_ctx *c, int field, int level, const float *host)
{
    if (!c || !host || field < 0 || field > 8 || level < 0 || level > 2) return fail(FDTD_EINVAL, "bad field/level");
    resolve_scales(c);
    std::vector<char> buf;
    to_storage(c, host, c->N, field_factor(c, field), buf);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(field_dev(c, field, level), buf.data(), c->N * c->esz, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int fdtd_ctx_get_field(fdtd_ctx *c, int field, int level, float *host)
{
    if (!c || !host || field < 0 || field > 8 || level < 0 || level > 2) return fail(FDTD_EINVAL, "bad field/level");
    resolve_scales(c);
    std::vector<char> buf(c->N * c->esz);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(buf.data(), field_dev(c, field, level), c->N * c->esz, hipMemcpyDeviceToHost));
    from_storage(c, buf.data(), c->N, field_factor(c, field), host);
    return 0;
}

extern "C" int fdtd_ctx_field_ptr(fdtd_ctx *c, int field, int level, void **dev)
{
    if (!c || !dev || field < 0 || field > 8 || level < 0 || level > 2) return fail(FDTD_EINVAL, "bad field/level");
    *dev = field_dev(c, field, level);
    return 0;
}

extern "C" int fdtd_step_fused(fdtd_ctx *c, hipStream_t stream)
{
    if (!c) return fail(FDTD_EINVAL, "null context");
    resolve_scales(c);
    int rc = upload_lut(c);
    if (rc) return rc;
    const int n = c->n;
    const int t = n % 3, s1 = (n + 2) % 3, s2 = (n + 1) % 3;
    const int ho = c->hcur, hn = c->hcur ^ 1;

    StepArgs a;
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
    int idx = (c->ext.jz_index_mode == FDTD_JZ_HOLD3) ? (n / 3) * 3 : n;
    const int lim = c->ext.jz_limit > 0 ? std::min<int>(c->ext.jz_limit, (int)c->jz.size()) : (int)c->jz.size();
    const float jz = (idx >= 0 && idx < lim) ? c->jz[idx] : 0.0f;
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz;
    a.pitch = c->pitch;
    a.plane = c->plane;
    if (c->storage == kStoreF32) {
        // the reference's coefficients, in float (fdtd_calculations.cu:49-75, :92-94)
        a.jz = jz;
        a.chx = c->dt / (c->mu0 * c->dx);
        a.chy = c->dt / (c->mu0 * c->dy);
        a.chz = c->dt / (c->mu0 * c->dz);
        a.cdx = c->dt / c->dx;
        a.cdy = c->dt / c->dy;
        a.cdz = c->dt / c->dz;
    } else {
        // scaled units: H~ = eta0 H / sH, D~ = D / (eps0 sD), E~ = E / sE, so
        // both curls carry c0 dt / h (eta0 / mu0 = 1 / (eps0 eta0) = c0)
        const double c0 = 1.0 / std::sqrt((double)c->mu0 * (double)c->eps0);
        a.jz = (float)((double)jz / ((double)c->eps0 * c->sD));
        a.chx = (float)(c0 * c->dt / c->dx * c->sE / c->sH);
        a.chy = (float)(c0 * c->dt / c->dy * c->sE / c->sH);
        a.chz = (float)(c0 * c->dt / c->dz * c->sE / c->sH);
        a.cdx = (float)(c0 * c->dt / c->dx * c->sH / c->sD);
        a.cdy = (float)(c0 * c->dt / c->dy * c->sH / c->sD);
        a.cdz = (float)(c0 * c->dt / c->dz * c->sH / c->sD);
    }
    a.kchunk = c->kchunk;
    a.tiles_x = c->tiles_x; a.tiles_y = c->tiles_y; a.tiles_z = c->tiles_z;
    a.xcd_remap = c->ext.xcd_remap;
    a.guard = (unsigned)c->guard;
    const bool timed = c->tused < c->tcap;
    if (timed) HIP_TRY(hipEventRecord(c->tev[2 * c->tused], stream));
    hipError_t e = launch_fused_step(a, c->ty, c->storage, stream);
    if (e != hipSuccess) return hip_fail(e, "fused_step");
    if (timed) {
        HIP_TRY(hipEventRecord(c->tev[2 * c->tused + 1], stream));
        c->tused++;
    }

    MurArgs m;
    m.exT = a.exT; m.eyT = a.eyT; m.ezT = a.ezT;
    m.exS = a.ex0; m.eyS = a.ey0; m.ezS = a.ez0;
    m.rpx0 = c->rp[0]; m.rpy0 = c->rp[1]; m.rpz0 = c->rp[2];
    m.rpxEnd = c->rp[3]; m.rpyEnd = c->rp[4]; m.rpzEnd = c->rp[5];
    m.nx = c->nx; m.ny = c->ny; m.nz = c->nz; m.pitch = c->pitch; m.plane = c->plane;
    m.dt = c->dt; m.hx = c->dx; m.hy = c->dy; m.hz = c->dz; m.mu0 = c->mu0; m.eps0 = c->eps0;
    mur_segments(c->nx, c->ny, c->nz, m.seg_start);
    e = launch_mur(m, c->storage, stream);
    if (e != hipSuccess) return hip_fail(e, "mur_faces");

    c->n = n + 1;
    c->hcur = hn;

    if (c->nprobe > 0 && c->probe_rows < c->probe_cap) {
        ProbeArgs p;
        for (int f = 0; f < 9; f++) {
            p.field[f] = field_dev(c, f, 0);
            p.factor[f] = field_factor(c, f);
        }
        p.probes = c->probes;
        p.nprobe = c->nprobe;
        p.pitch = c->pitch;
        p.plane = c->plane;
        p.out = c->probe_out;
        p.row = c->probe_rows++;
        e = launch_probes(p, c->storage, stream);
        if (e != hipSuccess) return hip_fail(e, "probe_cells");
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

extern "C" int fdtd_run(fdtd_ctx *c, int steps, hipStream_t stream)
{
    if (!c || steps < 0) return fail(FDTD_EINVAL, "bad run");
    for (int s = 0; s < steps; s++) {
        int rc = fdtd_step_fused(c, stream);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int fdtd_ctx_steps_done(fdtd_ctx *c) { return c ? c->n : FDTD_EINVAL; }

extern "C" int fdtd_probe_read(fdtd_ctx *c, double *host, int max_rows)
{
    if (!c || (!host && max_rows > 0)) return fail(FDTD_EINVAL, "bad probe read");
    HIP_TRY(hipDeviceSynchronize());
    const int rows = std::min(max_rows, c->probe_rows);
    if (rows > 0 && c->nprobe > 0)
        HIP_TRY(hipMemcpy(host, c->probe_out, sizeof(double) * (size_t)rows * c->nprobe, hipMemcpyDeviceToHost));
    return c->nprobe > 0 ? rows : 0;
}

extern "C" int fdtd_slices_read(fdtd_ctx *c, int *steps, float *data, int max_slices)
{
    if (!c) return fail(FDTD_EINVAL, "null context");
    if (c->side) HIP_TRY(hipStreamSynchronize(c->side));
    const int cnt = std::min(max_slices, c->slice_count);
    const size_t pl = (size_t)c->plane;
    for (int s = 0; s < cnt; s++) {
        if (steps) steps[s] = c->slice_steps[s];
        if (data)
            for (int f = 0; f < 9; f++)
                from_storage(c, c->slice_host + ((size_t)s * 9 + f) * pl * c->esz, pl, field_factor(c, f),
                             data + ((size_t)s * 9 + f) * pl);
    }
    return cnt;
}

extern "C" int fdtd_ctx_launch_shape(fdtd_ctx *c, int *tile_x, int *tile_y, int *kchunk, int *blocks)
{
    if (!c) return fail(FDTD_EINVAL, "null context");
    if (tile_x) *tile_x = 64;
    if (tile_y) *tile_y = c->ty;
    if (kchunk) *kchunk = c->kchunk;
    if (blocks) *blocks = c->tiles_x * c->tiles_y * c->tiles_z;
    return 0;
}

extern "C" int fdtd_ctx_set_timing(fdtd_ctx *c, int capacity)
{
    if (!c || capacity < 0) return fail(FDTD_EINVAL, "bad timing capacity");
    HIP_TRY(hipDeviceSynchronize());
    for (auto ev : c->tev) (void)hipEventDestroy(ev);
    c->tev.assign(2 * (size_t)capacity, nullptr);
    for (auto &ev : c->tev) HIP_TRY(hipEventCreate(&ev));
    c->tcap = capacity;
    c->tused = 0;
    return 0;
}

extern "C" int fdtd_ctx_kernel_time(fdtd_ctx *c, double *ms_total, int *launches)
{
    if (!c || !ms_total) return fail(FDTD_EINVAL, "null argument");
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

extern "C" int fdtd_ctx_scales(fdtd_ctx *c, double *scale_e, double *scale_d, double *scale_h)
{
    if (!c) return fail(FDTD_EINVAL, "null context");
    resolve_scales(c);
    if (scale_e) *scale_e = c->sE;
    if (scale_d) *scale_d = c->sD;
    if (scale_h) *scale_h = c->sH;
    return 0;
}
