// Host helpers shared by the single-grid context, the batch and the server:
// error reporting, 16-bit conversions and the per-step coefficients in every
// storage form.  One definition of each so that a grid stepped alone and
// the same grid stepped inside a batch see identical coefficients.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "fdtd_internal.h"

namespace {
thread_local std::string g_err;
}

int fdtd_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int fdtd_hip_fail(hipError_t e, const char *what)
{
    return fdtd_fail((int)e, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

extern "C" const char *fdtd_last_error(void) { return g_err.c_str(); }

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

void fdtd_scaled_coef(MatCoef &m, double dt, double eps0, double sg, double es, double ei, double tau, int debye_mode,
                      double ratio)
{
    // Scaled-unit E coefficients from the same a, b, c, d in double; ratio = sD / sE.
    const double a = 2.0 * eps0 * ei * tau;
    const double b = a + 2.0 * dt * (eps0 * es + sg * tau);
    const double c = 1.0 / (b + sg * dt * dt);
    const double d = (debye_mode == FDTD_DEBYE_REF ? 2.0 * b : a + b) - sg * dt * dt;
    m.k1 = (float)(c * d);
    m.k2 = (float)(c * a);
    m.k3 = (float)(c * eps0 * ratio * 2.0 * dt);
    m.k5 = (float)(c * eps0 * ratio * 2.0 * tau);
}

void fdtd_build_lut(const float *table, int nmat, float dt, float eps0, int debye_mode, int storage, double sE,
                    double sD, MatCoef *lut)
{
    for (int m = 0; m < nmat; m++) {
        const float *t = &table[4 * m];
        lut[m] = fdtd_mat_coef(dt, eps0, t[0], t[1], t[2], t[3], debye_mode);
        if (storage != kStoreF32) fdtd_scaled_coef(lut[m], dt, eps0, t[0], t[1], t[2], t[3], debye_mode, sD / sE);
    }
}

void fdtd_curl_coefs(const GridGeom &g, int storage, double sE, double sD, double sH, float ch[3], float cd[3])
{
    if (storage == kStoreF32) {
        // the reference's coefficients, in float (fdtd_calculations.cu:49-75, :92-94)
        ch[0] = g.dt / (g.mu0 * g.dx);
        ch[1] = g.dt / (g.mu0 * g.dy);
        ch[2] = g.dt / (g.mu0 * g.dz);
        cd[0] = g.dt / g.dx;
        cd[1] = g.dt / g.dy;
        cd[2] = g.dt / g.dz;
        return;
    }
    // scaled units: H~ = eta0 H / sH, D~ = D / (eps0 sD), E~ = E / sE, so both
    // curls carry c0 dt / h (eta0 / mu0 = 1 / (eps0 eta0) = c0)
    const double c0 = 1.0 / std::sqrt((double)g.mu0 * (double)g.eps0);
    const double h[3] = {g.dx, g.dy, g.dz};
    for (int q = 0; q < 3; q++) {
        ch[q] = (float)(c0 * g.dt / h[q] * sE / sH);
        cd[q] = (float)(c0 * g.dt / h[q] * sH / sD);
    }
}

float fdtd_jz_stored(float jz, int storage, float eps0, double sD)
{
    if (storage == kStoreF32) return jz;
    return (float)((double)jz / ((double)eps0 * sD));
}

float fdtd_jz_at(const float *jz, int jz_len, int jz_limit, int jz_mode, int n)
{
    const int idx = (jz_mode == FDTD_JZ_HOLD3) ? (n / 3) * 3 : n;
    const int lim = jz_limit > 0 ? (jz_limit < jz_len ? jz_limit : jz_len) : jz_len;
    return (idx >= 0 && idx < lim) ? jz[idx] : 0.0f;
}

void fdtd_auto_scales(const float *jz, int jz_len, float eps0, const FdtdExt &e, double *sE, double *sD, double *sH)
{
    // Peak fields a soft Jz source drives, measured against the FP64 oracle
    // in vacuum and in a dielectric (point, 10-cell line and full-plane
    // sources, 0.3-6 GHz pulses): |E| and |D|/eps0 stay below a third of
    // the accumulated kick K = max_t |sum_{s<=t} jz_s| / eps0 (a point
    // source reaches K/3 at its own cell), and eta0 |H| below 1.25 times the
    // largest single kick max |jz| / eps0.  Each bound maps to 4096 storage
    // units: 16x below the fp16 overflow (65504), with full fp16 precision
    // down to 1.5e-8 of the bound (the smallest normal, 6.1e-5 units).
    double peak = 0.0, acc = 0.0, accmax = 0.0;
    for (int t = 0; t < jz_len; t++) {
        peak = std::fmax(peak, (double)std::fabs(jz[t]));
        acc += (double)jz[t];
        accmax = std::fmax(accmax, std::fabs(acc));
    }
    constexpr double kUnits = 4096.0;
    const double se = accmax > 0.0 ? accmax / (double)eps0 / 3.0 / kUnits : 1.0;
    const double sh = peak > 0.0 ? 1.25 * peak / (double)eps0 / kUnits : 1.0;
    *sE = e.scale_e > 0.0f ? e.scale_e : se;
    *sD = e.scale_d > 0.0f ? e.scale_d : se;
    *sH = e.scale_h > 0.0f ? e.scale_h : sh;
}

double fdtd_field_factor(int storage, float mu0, float eps0, double sE, double sD, double sH, int field)
{
    if (storage == kStoreF32) return 1.0;
    const double eta0 = std::sqrt((double)mu0 / (double)eps0);
    if (field < 3) return (double)eps0 * sD;
    if (field < 6) return sH / eta0;
    return sE;
}

size_t fdtd_storage_bytes(int storage) { return (storage == kStoreF32 || storage == kStoreF32Fast) ? 4 : 2; }

void fdtd_to_storage(int storage, const float *in, size_t n, double factor, void *out)
{
    if (storage == kStoreF32) {
        memcpy(out, in, n * 4);
        return;
    }
    if (storage == kStoreF32Fast) {
        float *o = (float *)out;
        for (size_t t = 0; t < n; t++) o[t] = (float)((double)in[t] / factor);
        return;
    }
    uint16_t *o = (uint16_t *)out;
    for (size_t t = 0; t < n; t++) {
        const float v = (float)((double)in[t] / factor);
        o[t] = storage == kStoreF16 ? f32_to_f16(v) : f32_to_bf16(v);
    }
}

void fdtd_from_storage(int storage, const void *in, size_t n, double factor, float *out)
{
    if (storage == kStoreF32) {
        memcpy(out, in, n * 4);
        return;
    }
    if (storage == kStoreF32Fast) {
        const float *f = (const float *)in;
        for (size_t t = 0; t < n; t++) out[t] = (float)((double)f[t] * factor);
        return;
    }
    const uint16_t *h = (const uint16_t *)in;
    for (size_t t = 0; t < n; t++) {
        const float v = storage == kStoreF16 ? f16_to_f32(h[t]) : bf16_to_f32(h[t]);
        out[t] = (float)((double)v * factor);
    }
}

int fdtd_check_ext(const FdtdExt &e)
{
    if (e.debye_mode != FDTD_DEBYE_REF && e.debye_mode != FDTD_DEBYE_FIXED)
        return fdtd_fail(FDTD_EINVAL, "debye_mode %d", e.debye_mode);
    if (e.jz_index_mode != FDTD_JZ_STEP && e.jz_index_mode != FDTD_JZ_HOLD3)
        return fdtd_fail(FDTD_EINVAL, "jz_index_mode %d", e.jz_index_mode);
    if (e.tile_y != 0 && e.tile_y != 2 && e.tile_y != 4 && e.tile_y != 8) return fdtd_fail(FDTD_EINVAL, "tile_y %d", e.tile_y);
    if (e.kchunk < 0 || e.jz_limit < 0) return fdtd_fail(FDTD_EINVAL, "negative kchunk/jz_limit");
    if (e.xcd_remap < 0 || e.xcd_remap > 4) return fdtd_fail(FDTD_EINVAL, "xcd_remap %d", e.xcd_remap);
    if (e.storage != FDTD_STORE_F32 && e.storage != FDTD_STORE_F16 && e.storage != FDTD_STORE_BF16 &&
        e.storage != FDTD_STORE_F32_FAST)
        return fdtd_fail(FDTD_EINVAL, "storage %d", e.storage);
    if (!(e.scale_e >= 0.0f && e.scale_d >= 0.0f && e.scale_h >= 0.0f)) return fdtd_fail(FDTD_EINVAL, "negative scale");
    if (e.lane_cells < 0 || e.lane_cells > 2) return fdtd_fail(FDTD_EINVAL, "lane_cells %d", e.lane_cells);
    if (e.tile_x != 0 && e.tile_x != 64 && e.tile_x != 128) return fdtd_fail(FDTD_EINVAL, "tile_x %d", e.tile_x);
    if (e.load_zero_terms != 0 && e.load_zero_terms != 1)
        return fdtd_fail(FDTD_EINVAL, "load_zero_terms %d", e.load_zero_terms);
    if (e.time_steps < 0 || e.time_steps > 2) return fdtd_fail(FDTD_EINVAL, "time_steps %d", e.time_steps);
    return 0;
}

int fdtd_check_sources(int nx, int ny, int nz, int nsrc, const int *xyz)
{
    for (int s = 0; s < nsrc; s++) {
        const int x = xyz[3 * s], y = xyz[3 * s + 1], z = xyz[3 * s + 2];
        // updateSources reads Hy at x-1 and Hx at y-1 (fdtd_calculations.cu:205-208), so x and y
        // start at 1; the upper faces and the top plane are allowed, as the reference writes Dz there.
        if (x < 1 || x >= nx || y < 1 || y >= ny || z < 0 || z >= nz)
            return fdtd_fail(FDTD_ERANGE, "source %d at (%d,%d,%d) outside 1<=x<nx, 1<=y<ny, 0<=z<nz", s, x, y, z);
    }
    return 0;
}

int fdtd_check_probes(int nx, int ny, int nz, int nprobe, const int *probes)
{
    for (int p = 0; p < nprobe; p++) {
        const int *q = probes + 4 * p;
        if (q[0] < 0 || q[0] > 8 || q[1] < 0 || q[1] >= nx || q[2] < 0 || q[2] >= ny || q[3] < 0 || q[3] >= nz)
            return fdtd_fail(FDTD_ERANGE, "probe %d out of range", p);
    }
    return 0;
}

int fdtd_probe_peaks(int storage, const double *series, int rows, int nprobe, const int *fields, const double *factor,
                     double *peak)
{
    const double f16_min_normal = 6.103515625e-05;   // 2^-14
    int low = 0;
    for (int p = 0; p < nprobe; p++) {
        double m = 0.0;
        for (int r = 0; r < rows; r++) m = std::max(m, std::fabs(series[(size_t)r * nprobe + p]));
        const double f = std::fabs(factor[fields[p]]);
        peak[p] = f > 0.0 ? m / f : m;
        if (storage == kStoreF16 && peak[p] > 0.0 && peak[p] < f16_min_normal) low++;
    }
    return low;
}

void fdtd_to_storage_pitched(int storage, const float *in, int nx, size_t rows, int pitch, double factor, void *out)
{
    const size_t esz = fdtd_storage_bytes(storage);
    char *o = (char *)out;
    for (size_t r = 0; r < rows; r++) {
        char *row = o + r * pitch * esz;
        fdtd_to_storage(storage, in + r * nx, nx, factor, row);
        memset(row + (size_t)nx * esz, 0, (size_t)(pitch - nx) * esz);
    }
}

void fdtd_from_storage_pitched(int storage, const void *in, int nx, size_t rows, int pitch, double factor, float *out)
{
    const size_t esz = fdtd_storage_bytes(storage);
    const char *i = (const char *)in;
    for (size_t r = 0; r < rows; r++) fdtd_from_storage(storage, i + r * pitch * esz, nx, factor, out + r * nx);
}

void fdtd_pitch_bytes(const unsigned char *in, int nx, size_t rows, int pitch, unsigned char *out)
{
    for (size_t r = 0; r < rows; r++) {
        memcpy(out + r * pitch, in + r * nx, nx);
        memset(out + r * pitch + nx, 0, pitch - nx);
    }
}

void mur_x_coefs(int ny, int nz, float dt, float hx, float mu0, float eps0, const float *rpx0, const float *rpxEnd,
                 float4 *out0, float4 *out1)
{
    auto one = [&](float rpf, float rpn) {
        const float sf = std::sqrt(mu0 * eps0 * rpf), sn = std::sqrt(mu0 * eps0 * rpn);
        float4 c;
        c.x = 1 / (dt + hx * sf);
        c.y = dt - hx * sn;
        c.z = dt + hx * sn;
        c.w = dt - hx * sf;
        return c;
    };
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            const size_t r = ((size_t)k * ny + j) * 2, q = (size_t)k * ny + j;
            out0[q] = one(rpx0[r], rpx0[r + 1]);
            out1[q] = one(rpxEnd[r + 1], rpxEnd[r]);
        }
}

extern "C" int fdtd_stream_probe_layout(long long cells, int row, int blocks, int reps, double *gbs,
                                        double *ms_per_pass)
{
    if (cells < 1 || reps < 1 || !gbs || row < 0)
        return fdtd_fail(FDTD_EINVAL, "stream probe: cells %lld, row %d, reps %d", cells, row, reps);
    if (row > 0 && cells % row != 0) return fdtd_fail(FDTD_EINVAL, "stream probe: cells %lld not whole rows of %d", cells, row);
    if (blocks <= 0) blocks = stream_mix_resident_blocks();
    if (blocks <= 0) return fdtd_fail(FDTD_EINVAL, "stream probe: no resident blocks");
    // separate arrays: the contexts' stagger between arrays; row records: one block
    const size_t stride = row > 0 ? (size_t)cells : (size_t)cells + 65536;
    float *blob = nullptr;
    HIP_TRY(hipMalloc(&blob, 24 * stride * sizeof(float)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    StreamMix p;
    for (int a = 0; a < 24; a++) {
        float *b = blob + (row > 0 ? 0 : a * stride);
        if (a < 15) p.r[a] = b;
        else p.w[a - 15] = b;
    }
    p.n = cells;
    p.row = row;
    hipError_t e = hipMemset(blob, 0, 24 * stride * sizeof(float));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_stream_mix(p, blocks, nullptr);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_stream_mix(p, blocks, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(blob);
    if (e != hipSuccess) return fdtd_hip_fail(e, "stream probe");
    const double per = (double)ms / reps;
    *gbs = 24.0 * sizeof(float) * (double)cells / (per * 1e-3) / 1e9;
    if (ms_per_pass) *ms_per_pass = per;
    return 0;
}

extern "C" int fdtd_stream_probe(long long cells, int blocks, int reps, double *gbs, double *ms_per_pass)
{
    return fdtd_stream_probe_layout(cells, 0, blocks, reps, gbs, ms_per_pass);
}
