/*
 * fdtd_oracle.c -- CPU restatement of the FDTD time-stepping path of
 * rafalgrodzinski/fdtd, used ONLY as test infrastructure.
 *
 * Who may use this library: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker or as the timed CPU baseline.
 * The product (fdtd_amd/) never links, loads or calls it.
 *
 * Two arithmetic levels:
 *   * FP32 transcription ("or_*_f32"): the reference update equations with
 *     the reference's operand order, ranges and the FP64 promotions that its
 *     double literals cause (fdtd_calculations.cu:141-181).  Built with
 *     -ffp-contract=off so no multiply-add is fused.
 *   * FP64 oracle ("or_*_f64"): the same equations entirely in double.
 * Debye mode: OR_DEBYE_REF (d = 2b - sigma dt^2, fdtd_calculations.cu:144,
 * unconditionally unstable) or OR_DEBYE_FIXED (d = a + b - sigma dt^2).
 *
 * What pins each function (tests/test_oracle_*.py):
 *   or_derive_params      closed forms (dx = lambda/epw, Courant dt) and the
 *                         values derived for data/input_params.
 *   or_setup_jz           known answers for data/input_params (fine, temp,
 *                         istart, sample count, support, peak) and the
 *                         telescoping-sum invariant of a central difference.
 *   or_setup_rp           closed form Re(eps(omega)) evaluated in numpy for
 *                         both the x-plane (no I) and y/z-plane forms.
 *   or_update_h_*         discrete div(H) = 0 (div curl E = 0) on random E,
 *                         mirror symmetry of a centred source.
 *   or_update_d_*         discrete div(D) = 0 away from sources and
 *                         div(D) = -sum(jz)/dz at a source (charge build-up).
 *   or_update_e_*         single-cell recurrences: REF growth factor = root of
 *                         r^2 - c d r + c a = 0 (numpy.roots), FIXED steady
 *                         state recovers eps(omega) of a Debye medium.
 *   or_update_mur_*       reduces in vacuum to the textbook first-order Mur
 *                         form E0' = E1 + (S-1)/(S+1)(E1' - E0); with rp
 *                         varying cell by cell, the residual of the one-way
 *                         wave equation with each node at its own speed
 *                         (fixes which rp plane is the face's, p0, and which
 *                         the neighbour's, p1); absorbed energy vs a closed
 *                         box.
 *   or_update_sources_*   linearity in jz and the 3-step sample hold.
 *   or_load_specs/pgm     the tissue table fixture (53 rows) and a PGM file in
 *                         the reference's layout.
 * Parity unpinned: nvcc's exact FMA contraction on the reference device code
 * (no NVIDIA GPU or nvcc exists here); where the FP64 promotions sit in
 * or_update_e_f32 (as read in the reference's text: a moved promotion
 * changes results by ulps, inside every FP64 bound); the reference's
 * undefined behaviour
 * (out-of-range coefficient loads, specs[94], dJz past 512) is replaced by
 * defined behaviour: no out-of-range loads, code 94 stored as parsed, and
 * jz = 0 past the end of the table.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_DEBYE_REF 0
#define OR_DEBYE_FIXED 1

#define IDX(i, j, k) ((size_t)(k) * (size_t)ny * (size_t)nx + (size_t)(j) * (size_t)nx + (size_t)(i))

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int or_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Parameters: fdtd.cu:547-637                                        */
/* ------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz;
    int iterations;
    int epw;
    float wave_frequency, pulse_width, pulse_modulation_frequency;
    int nsrc;
    int src[3 * 64];
    float sigma, eps_s, eps_i, tau_d;
    float pi, c, timeskip, lambda, dt, dx, dy, dz, mu0, eps0;
    char input_path[1024], output_path[1024];
} or_params;

/* Derivations of fdtd.cu:625-634: the float fields hold the results of
 * double (or float) expressions exactly as C evaluates them there. */
void or_derive_params(or_params *p, double timeskip)
{
    p->pi = acos(-1.0);
    p->c = 3.0 * pow(10.0, 8.0);
    p->timeskip = (float)timeskip;
    p->lambda = p->c / p->wave_frequency;           /* float / float */
    p->dx = p->lambda / p->epw;                     /* float / int */
    p->dy = p->dx;
    p->dz = p->dx;
    p->dt = 1.0 * p->timeskip / (p->c * sqrt(1.0 / pow(p->dx, 2.0) + 1.0 / pow(p->dy, 2.0) + 1.0 / pow(p->dz, 2.0)));
    p->mu0 = 4.0 * p->pi * pow(10.0, -7.0);
    p->eps0 = 1.0 / (p->mu0 * p->c * p->c);         /* (float*float*float) in float, then 1.0/ in double */
}

/* Positional parse of fdtd.cu:559-620 (each line is "<key> <values>";
 * unused lines are skipped whole). t_max is rounded up to a multiple of 3
 * (fdtd.cu:563) and source coordinates become 0-based (fdtd.cu:607-609).
 * Returns 0 on success, -1 if the file cannot be opened, -2 on a short file. */
int or_parse_params(const char *path, or_params *p)
{
    char t[1024];
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    memset(p, 0, sizeof(*p));
    int ok = 1;
    ok &= fscanf(f, "%s %d %d %d\n", t, &p->nx, &p->ny, &p->nz) == 4;
    ok &= fscanf(f, "%s %d\n", t, &p->iterations) == 2;
    p->iterations = ((p->iterations - 1) / 3 + 1) * 3;
    ok &= fgets(t, sizeof t, f) != NULL;
    ok &= fscanf(f, "%s %1023s\n", t, p->input_path) == 2;
    ok &= fgets(t, sizeof t, f) != NULL;
    ok &= fscanf(f, "%s %1023s\n", t, p->output_path) == 2;
    for (int u = 0; u < 9; u++) ok &= fgets(t, sizeof t, f) != NULL;
    ok &= fscanf(f, "%s %d\n", t, &p->epw) == 2;
    ok &= fscanf(f, "%s %g\n", t, &p->wave_frequency) == 2;
    ok &= fscanf(f, "%s %g\n", t, &p->pulse_width) == 2;
    ok &= fscanf(f, "%s %g\n", t, &p->pulse_modulation_frequency) == 2;
    ok &= fscanf(f, "%s %d\n", t, &p->nsrc) == 2;
    if (!ok || p->nsrc < 0 || p->nsrc > 64) { fclose(f); return -2; }
    for (int s = 0; s < p->nsrc; s++) {
        ok &= fscanf(f, "%s %d %d %d\n", t, &p->src[3 * s], &p->src[3 * s + 1], &p->src[3 * s + 2]) == 4;
        p->src[3 * s] -= 1; p->src[3 * s + 1] -= 1; p->src[3 * s + 2] -= 1;
    }
    ok &= fgets(t, sizeof t, f) != NULL;
    ok &= fscanf(f, "%s %f\n", t, &p->sigma) == 2;
    ok &= fscanf(f, "%s %f\n", t, &p->eps_s) == 2;
    ok &= fscanf(f, "%s %f\n", t, &p->eps_i) == 2;
    ok &= fscanf(f, "%s %f\n", t, &p->tau_d) == 2;
    fclose(f);
    if (!ok) return -2;
    or_derive_params(p, 1.0);
    return 0;
}

size_t or_params_size(void) { return sizeof(or_params); }

/* ------------------------------------------------------------------ */
/* Source waveform: fdtd.cu:1027-1082                                 */
/* ------------------------------------------------------------------ */

/* Fills jz[65536]. info[0..3] = fine, temp, istart, raw sample count.
 * Returns 0, or -1 if fine == 0 (the reference divides by it), -2 if no
 * start sample exists (the reference then reads an uninitialised istart),
 * -3 if the raw samples would overflow the table. */
int or_setup_jz(float pulse_width, float wave_frequency, float pmf, float dt, float dx, float dy, float dz,
                float *jz, int *info)
{
    const int off = 1 << 16;
    float *tmp = (float *)calloc((size_t)off * 2, sizeof(float));
    float *tmp2 = (float *)calloc((size_t)off * 2, sizeof(float));
    memset(jz, 0, sizeof(float) * off);
    int fine = (1 << 13) * pulse_width * wave_frequency * dt;                 /* float chain, truncated */
    if (fine == 0) { free(tmp); free(tmp2); return -1; }
    int temp = 1.0 / (pulse_width * wave_frequency) / (dt / fine) / 2.0;      /* double chain, truncated */
    if (temp + 1 >= off) { free(tmp); free(tmp2); return -3; }
    for (int i = -temp - 1; i <= temp + 1; i++) {
        float v1 = ((float)i / (((float)temp + 1.0) / 4.0));
        float v2 = exp(-pow(v1, 2.0));
        float v3 = cos(2.0 * acos(-1.0) * pmf * wave_frequency * i * (dt / fine));
        tmp[i + off] = v2 * v3;
    }
    int istart = 0, found = 0;
    for (int i = -(1 << 12); i < (1 << 12); i++) {
        if ((fabs(tmp[i + off]) >= 1e-9) && (i % fine == 0)) { istart = i; found = 1; break; }
    }
    if (!found) { free(tmp); free(tmp2); return -2; }
    int n = 0;
    for (int i = istart; i <= temp + 1; i += fine) {
        if (n >= off) { free(tmp); free(tmp2); return -3; }
        jz[n++] = tmp[i + off] * 1e-15 / dt / 3.0;
    }
    for (int i = 2; i <= (1 << 14); i++) {
        float v = (((jz[i] - jz[i - 1]) / dt) + ((jz[i - 1] - jz[i - 2]) / dt)) / 2.0 * (dt * dz) / (dx * dy * dz);
        tmp2[i - 1 + off] = v;
    }
    for (int i = 0; i < (1 << 14); i++) jz[i] = tmp2[i + off + 1];
    if (info) { info[0] = fine; info[1] = temp; info[2] = istart; info[3] = n; }
    free(tmp); free(tmp2);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Mur permittivity planes: fdtd.cu:942-1024                          */
/* ------------------------------------------------------------------ */

static float rp_x(float f, float pi, float eps0, float sg, float es, float ei, float tau)
{
    float complex c1 = 0.0 + 2.0 * pi * f * tau;
    float complex c2 = 0.0 + sg / (2.0 * pi * f * eps0);
    return creal(ei + (es - ei) / (1.0 + c1) - c2);
}

static float rp_yz(float f, float pi, float eps0, float sg, float es, float ei, float tau)
{
    float complex c1 = 0.0 + 2.0 * pi * f * tau * I;
    float complex c2 = 0.0 + sg / (2.0 * pi * f * eps0) * I;
    return creal(ei + (es - ei) / (1.0 + c1) - c2);
}

/* Material arrays are the 4 per-voxel float arrays of the reference.
 * rp planes: x planes 2*ny*nz (rpnx=2, rpny=ny), y planes nx*2*nz, z planes nx*ny*2. */
void or_setup_rp(int nx, int ny, int nz, float f, float pi, float eps0,
                 const float *sigma, const float *epsS, const float *epsI, const float *tau,
                 float *rpx0, float *rpy0, float *rpz0, float *rpxEnd, float *rpyEnd, float *rpzEnd)
{
#define M4(i, j, k) sigma[IDX(i, j, k)], epsS[IDX(i, j, k)], epsI[IDX(i, j, k)], tau[IDX(i, j, k)]
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            for (int i = 0; i < 2; i++) rpx0[(size_t)k * ny * 2 + j * 2 + i] = rp_x(f, pi, eps0, M4(i, j, k));
            for (int i = nx - 2; i < nx; i++) rpxEnd[(size_t)k * ny * 2 + j * 2 + (i - (nx - 2))] = rp_x(f, pi, eps0, M4(i, j, k));
        }
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
            for (int j = 0; j < 2; j++) rpy0[(size_t)k * 2 * nx + j * nx + i] = rp_yz(f, pi, eps0, M4(i, j, k));
            for (int j = ny - 2; j < ny; j++) rpyEnd[(size_t)k * 2 * nx + (j - (ny - 2)) * nx + i] = rp_yz(f, pi, eps0, M4(i, j, k));
        }
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            for (int k = 0; k < 2; k++) rpz0[(size_t)k * ny * nx + j * nx + i] = rp_yz(f, pi, eps0, M4(i, j, k));
            for (int k = nz - 2; k < nz; k++) rpzEnd[(size_t)(k - (nz - 2)) * ny * nx + j * nx + i] = rp_yz(f, pi, eps0, M4(i, j, k));
        }
#undef M4
}

/* ------------------------------------------------------------------ */
/* Tissue table and PGM slices: fdtd.cu:875-939                       */
/* ------------------------------------------------------------------ */

/* specs[95*4] = {sigma, eps_s, eps_i, tau} per code; rows are read until a
 * code >= 94 (that row is stored, as parsed) or 94 rows.  Returns rows read or -1. */
int or_load_specs(const char *path, float *specs)
{
    char t[1024];
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    memset(specs, 0, sizeof(float) * 95 * 4);
    int rows = 0;
    for (int r = 0; r < 94; r++) {
        int code; float s, es, ei, tau;
        if (fscanf(f, "%d %s %g %g %g %g\n", &code, t, &s, &es, &ei, &tau) != 6) break;
        if (code < 0 || code > 94) { fclose(f); return -2; }
        specs[code * 4 + 0] = s; specs[code * 4 + 1] = es; specs[code * 4 + 2] = ei; specs[code * 4 + 3] = tau;
        rows++;
        if (code >= 94) break;
    }
    fclose(f);
    return rows;
}

/* Reads the first n integers after a 6-token header (width ignored, as the
 * reference does at fdtd.cu:920-926).  Returns 0, -1 open failure, -2 short. */
int or_load_pgm_codes(const char *path, int n, int *codes, int *width, int *height)
{
    char t[1024];
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    int w = 0, h = 0;
    if (fscanf(f, "%s %s %s %d %d %s", t, t, t, &w, &h, t) != 6) { fclose(f); return -2; }
    for (int q = 0; q < n; q++)
        if (fscanf(f, "%d", &codes[q]) != 1) { fclose(f); return -2; }
    fclose(f);
    if (width) *width = w;
    if (height) *height = h;
    return 0;
}

/* ------------------------------------------------------------------ */
/* FP32 transcription of the device kernels: fdtd_calculations.cu      */
/* ------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz;
    float dt, dx, dy, dz, mu0, eps0;
} or_grid;

/* fdtd_calculations.cu:32-77 (in place; the second Hx term uses dy). */
void or_update_h_f32(const or_grid *g, float *hx, float *hy, float *hz,
                     const float *ex, const float *ey, const float *ez)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const float dt = g->dt, dx = g->dx, dy = g->dy, dz = g->dz, mu0 = g->mu0;
    const size_t sy = nx, sz = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = IDX(i, j, k);
                if (i >= 1 && i < nx - 1 && j < ny - 1 && k < nz - 1)
                    hx[o] = hx[o] - dt / (mu0 * dy) * (ez[o + sy] - ez[o]) + dt / (mu0 * dy) * (ey[o + sz] - ey[o]);
                if (i < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1)
                    hy[o] = hy[o] - dt / (mu0 * dz) * (ex[o + sz] - ex[o]) + dt / (mu0 * dx) * (ez[o + 1] - ez[o]);
                if (i < nx - 1 && j < ny - 1 && k >= 1 && k < nz - 1)
                    hz[o] = hz[o] - dt / (mu0 * dx) * (ey[o + 1] - ey[o]) + dt / (mu0 * dy) * (ex[o + sy] - ex[o]);
            }
}

/* fdtd_calculations.cu:80-122 (out of place; cells outside the ranges keep
 * the target's value). */
void or_update_d_f32(const or_grid *g, float *dxT, float *dyT, float *dzT,
                     const float *dxS, const float *dyS, const float *dzS,
                     const float *hx, const float *hy, const float *hz)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const float x = g->dt / g->dx, y = g->dt / g->dy, z = g->dt / g->dz;
    const size_t sy = nx, sz = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = IDX(i, j, k);
                if (i < nx - 1 && j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1)
                    dxT[o] = dxS[o] + y * (hz[o] - hz[o - sy]) - z * (hy[o] - hy[o - sz]);
                if (i >= 1 && i < nx - 1 && j < ny - 1 && k >= 1 && k < nz - 1)
                    dyT[o] = dyS[o] + z * (hx[o] - hx[o - sz]) - x * (hz[o] - hz[o - 1]);
                if (i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1)
                    dzT[o] = dzS[o] + x * (hy[o] - hy[o - 1]) - y * (hx[o] - hx[o - sy]);
            }
}

/* fdtd_calculations.cu:187-213: every listed source cell gets the D update
 * recomputed from the source slot minus jz.  Sources need x >= 1 and y >= 1
 * (the reference reads x-1 and y-1). */
void or_update_sources_f32(const or_grid *g, float *dzT, const float *dzS, const float *hx, const float *hy,
                           int nsrc, const int *src, float jz)
{
    const int nx = g->nx, ny = g->ny;
    for (int s = 0; s < nsrc; s++) {
        int x = src[3 * s], y = src[3 * s + 1], z = src[3 * s + 2];
        size_t o = IDX(x, y, z);
        float val = dzS[o] + g->dt / g->dx * (hy[o] - hy[o - 1]) - g->dt / g->dy * (hx[o] - hx[o - (size_t)nx]) - jz;
        dzT[o] = val;
    }
}

/* fdtd_calculations.cu:125-184.  a, b, c, d are double expressions rounded
 * to float; the D terms and the sum are double; the result is rounded once. */
void or_update_e_f32(const or_grid *g, int debye_mode,
                     float *exT, float *eyT, float *ezT,
                     const float *ex0, const float *ey0, const float *ez0,
                     const float *ex1, const float *ey1, const float *ez1,
                     const float *dx0, const float *dy0, const float *dz0,
                     const float *dx1, const float *dy1, const float *dz1,
                     const float *dx2, const float *dy2, const float *dz2,
                     const float *sigma, const float *epsI, const float *epsS, const float *tauD)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const float dDt = g->dt, dEps0 = g->eps0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = IDX(i, j, k);
                float sg = sigma[o], ei = epsI[o], es = epsS[o], tau = tauD[o];
                float a = 2.0 * dEps0 * ei * tau;
                float b = a + 2.0 * dDt * (dEps0 * es + sg * tau);
                float c = 1.0 / (b + sg * dDt * dDt);
                float d = (debye_mode == OR_DEBYE_REF) ? (float)(2.0 * b - sg * dDt * dDt)
                                                       : (float)((double)a + (double)b - (double)(sg * dDt * dDt));
                if (i < nx - 1 && j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1)
                    exT[o] = c * (d * ex0[o] - a * ex1[o] + (2.0 * (dDt + tau)) * dx0[o]
                                  - (2.0 * dDt + 4.0 * tau) * dx1[o] + (2.0 * tau) * dx2[o]);
                if (i >= 1 && i <= nx - 1 && j <= ny - 1 && k >= 1 && k <= nz - 1) {
                    eyT[o] = c * (d * ey0[o] - a * ey1[o] + (2.0 * (dDt + tau)) * dy0[o]
                                  - (2.0 * dDt + 4.0 * tau) * dy1[o] + (2.0 * tau) * dy2[o]);
                    ezT[o] = c * (d * ez0[o] - a * ez1[o] + (2.0 * (dDt + tau)) * dz0[o]
                                  - (2.0 * dDt + 4.0 * tau) * dz1[o] + (2.0 * tau) * dz2[o]);
                }
            }
}

/* One first-order Mur face value (fdtd_calculations.cu:240-248 pattern):
 * p0 = face rp, p1 = neighbour rp, h = cell size along the normal. */
static inline float mur_f32(float dt, float h, float mu0, float eps0, float rpf, float rpn,
                            float tgt_nb, float src_nb, float src_f)
{
    return 1 / (dt + h * sqrtf(mu0 * eps0 * rpf)) *
           ((dt - h * sqrtf(mu0 * eps0 * rpn)) * tgt_nb +
            (dt + h * sqrtf(mu0 * eps0 * rpn)) * src_nb -
            (dt - h * sqrtf(mu0 * eps0 * rpf)) * src_f);
}

/* fdtd_calculations.cu:216-428.  The Ey x-high face uses the source
 * neighbour twice (fdtd_calculations.cu:323-325). */
void or_update_mur_f32(const or_grid *g, float *exT, float *eyT, float *ezT,
                       const float *exS, const float *eyS, const float *ezS,
                       const float *rpx0, const float *rpy0, const float *rpz0,
                       const float *rpxEnd, const float *rpyEnd, const float *rpzEnd)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const float dt = g->dt, hx = g->dx, hy = g->dy, hz = g->dz, m = g->mu0, e = g->eps0;
    const size_t sy = nx, sz = (size_t)nx * ny;
#define RPY(p, i, j, k) p[(size_t)(k) * 2 * nx + (size_t)(j) * nx + (i)]
#define RPZ(p, i, j, k) p[(size_t)(k) * ny * nx + (size_t)(j) * nx + (i)]
#define RPX(p, i, j, k) p[(size_t)(k) * ny * 2 + (size_t)(j) * 2 + (i)]
    /* Ex: y faces, then z faces */
    for (int k = 1; k < nz - 1; k++)
        for (int i = 0; i < nx - 1; i++) {
            size_t o = IDX(i, 0, k);
            exT[o] = mur_f32(dt, hy, m, e, RPY(rpy0, i, 0, k), RPY(rpy0, i, 1, k), exT[o + sy], exS[o + sy], exS[o]);
            o = IDX(i, ny - 1, k);
            exT[o] = mur_f32(dt, hy, m, e, RPY(rpyEnd, i, 1, k), RPY(rpyEnd, i, 0, k), exT[o - sy], exS[o - sy], exS[o]);
        }
    for (int j = 1; j < ny - 1; j++)
        for (int i = 0; i < nx - 1; i++) {
            size_t o = IDX(i, j, 0);
            exT[o] = mur_f32(dt, hz, m, e, RPZ(rpz0, i, j, 0), RPZ(rpz0, i, j, 1), exT[o + sz], exS[o + sz], exS[o]);
            o = IDX(i, j, nz - 1);
            exT[o] = mur_f32(dt, hz, m, e, RPZ(rpzEnd, i, j, 1), RPZ(rpzEnd, i, j, 0), exT[o - sz], exS[o - sz], exS[o]);
        }
    /* Ey: x faces (x-high quirk), then z faces */
    for (int k = 1; k < nz - 1; k++)
        for (int j = 0; j < ny - 1; j++) {
            size_t o = IDX(0, j, k);
            eyT[o] = mur_f32(dt, hx, m, e, RPX(rpx0, 0, j, k), RPX(rpx0, 1, j, k), eyT[o + 1], eyS[o + 1], eyS[o]);
            o = IDX(nx - 1, j, k);
            eyT[o] = mur_f32(dt, hx, m, e, RPX(rpxEnd, 1, j, k), RPX(rpxEnd, 0, j, k), eyS[o - 1], eyS[o - 1], eyS[o]);
        }
    for (int j = 0; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            size_t o = IDX(i, j, 0);
            eyT[o] = mur_f32(dt, hz, m, e, RPZ(rpz0, i, j, 0), RPZ(rpz0, i, j, 1), eyT[o + sz], eyS[o + sz], eyS[o]);
            o = IDX(i, j, nz - 1);
            eyT[o] = mur_f32(dt, hz, m, e, RPZ(rpzEnd, i, j, 1), RPZ(rpzEnd, i, j, 0), eyT[o - sz], eyS[o - sz], eyS[o]);
        }
    /* Ez: x faces, then y faces */
    for (int k = 0; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++) {
            size_t o = IDX(0, j, k);
            ezT[o] = mur_f32(dt, hx, m, e, RPX(rpx0, 0, j, k), RPX(rpx0, 1, j, k), ezT[o + 1], ezS[o + 1], ezS[o]);
            o = IDX(nx - 1, j, k);
            ezT[o] = mur_f32(dt, hx, m, e, RPX(rpxEnd, 1, j, k), RPX(rpxEnd, 0, j, k), ezT[o - 1], ezS[o - 1], ezS[o]);
        }
    for (int k = 0; k < nz - 1; k++)
        for (int i = 1; i < nx - 1; i++) {
            size_t o = IDX(i, 0, k);
            ezT[o] = mur_f32(dt, hy, m, e, RPY(rpy0, i, 0, k), RPY(rpy0, i, 1, k), ezT[o + sy], ezS[o + sy], ezS[o]);
            o = IDX(i, ny - 1, k);
            ezT[o] = mur_f32(dt, hy, m, e, RPY(rpyEnd, i, 1, k), RPY(rpyEnd, i, 0, k), ezT[o - sy], ezS[o - sy], ezS[o]);
        }
}

/* ------------------------------------------------------------------ */
/* FP64 oracle: the same equations, all arithmetic in double           */
/* ------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz;
    double dt, dx, dy, dz, mu0, eps0;
} or_grid64;

static void grid64(const or_grid *g, or_grid64 *o)
{
    o->nx = g->nx; o->ny = g->ny; o->nz = g->nz;
    o->dt = g->dt; o->dx = g->dx; o->dy = g->dy; o->dz = g->dz; o->mu0 = g->mu0; o->eps0 = g->eps0;
}

static void upd_h64(const or_grid64 *g, double *hx, double *hy, double *hz, const double *ex, const double *ey, const double *ez)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double cy = g->dt / (g->mu0 * g->dy), cz = g->dt / (g->mu0 * g->dz), cx = g->dt / (g->mu0 * g->dx);
    const size_t sy = nx, sz = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = IDX(i, j, k);
                if (i >= 1 && i < nx - 1 && j < ny - 1 && k < nz - 1)
                    hx[o] = hx[o] - cy * (ez[o + sy] - ez[o]) + cy * (ey[o + sz] - ey[o]);
                if (i < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1)
                    hy[o] = hy[o] - cz * (ex[o + sz] - ex[o]) + cx * (ez[o + 1] - ez[o]);
                if (i < nx - 1 && j < ny - 1 && k >= 1 && k < nz - 1)
                    hz[o] = hz[o] - cx * (ey[o + 1] - ey[o]) + cy * (ex[o + sy] - ex[o]);
            }
}

static void upd_d64(const or_grid64 *g, double *dxT, double *dyT, double *dzT, const double *dxS, const double *dyS,
                    const double *dzS, const double *hx, const double *hy, const double *hz)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double x = g->dt / g->dx, y = g->dt / g->dy, z = g->dt / g->dz;
    const size_t sy = nx, sz = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = IDX(i, j, k);
                if (i < nx - 1 && j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1)
                    dxT[o] = dxS[o] + y * (hz[o] - hz[o - sy]) - z * (hy[o] - hy[o - sz]);
                if (i >= 1 && i < nx - 1 && j < ny - 1 && k >= 1 && k < nz - 1)
                    dyT[o] = dyS[o] + z * (hx[o] - hx[o - sz]) - x * (hz[o] - hz[o - 1]);
                if (i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1)
                    dzT[o] = dzS[o] + x * (hy[o] - hy[o - 1]) - y * (hx[o] - hx[o - sy]);
            }
}

static void upd_src64(const or_grid64 *g, double *dzT, const double *dzS, const double *hx, const double *hy,
                      int nsrc, const int *src, double jz)
{
    const int nx = g->nx, ny = g->ny;
    for (int s = 0; s < nsrc; s++) {
        size_t o = IDX(src[3 * s], src[3 * s + 1], src[3 * s + 2]);
        dzT[o] = dzS[o] + g->dt / g->dx * (hy[o] - hy[o - 1]) - g->dt / g->dy * (hx[o] - hx[o - (size_t)nx]) - jz;
    }
}

static void upd_e64(const or_grid64 *g, int debye_mode, double *T[3], double *E0[3], double *E1[3], double *D0[3],
                    double *D1[3], double *D2[3], const float *sigma, const float *epsI, const float *epsS, const float *tauD)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double dt = g->dt, e0 = g->eps0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = IDX(i, j, k);
                double sg = sigma[o], ei = epsI[o], es = epsS[o], tau = tauD[o];
                double a = 2.0 * e0 * ei * tau;
                double b = a + 2.0 * dt * (e0 * es + sg * tau);
                double c = 1.0 / (b + sg * dt * dt);
                double d = (debye_mode == OR_DEBYE_REF) ? 2.0 * b - sg * dt * dt : a + b - sg * dt * dt;
                double g1 = 2.0 * (dt + tau), g2 = 2.0 * dt + 4.0 * tau, g3 = 2.0 * tau;
                int cx = (i < nx - 1 && j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1);
                int cyz = (i >= 1 && j <= ny - 1 && k >= 1);
                for (int q = 0; q < 3; q++) {
                    if (q == 0 ? !cx : !cyz) continue;
                    T[q][o] = c * (d * E0[q][o] - a * E1[q][o] + g1 * D0[q][o] - g2 * D1[q][o] + g3 * D2[q][o]);
                }
            }
}

static inline double mur64(double dt, double h, double mu0, double eps0, double rpf, double rpn, double tn, double sn, double sf)
{
    return 1.0 / (dt + h * sqrt(mu0 * eps0 * rpf)) *
           ((dt - h * sqrt(mu0 * eps0 * rpn)) * tn + (dt + h * sqrt(mu0 * eps0 * rpn)) * sn - (dt - h * sqrt(mu0 * eps0 * rpf)) * sf);
}

static void upd_mur64(const or_grid64 *g, double *exT, double *eyT, double *ezT, const double *exS, const double *eyS,
                      const double *ezS, const float *rpx0, const float *rpy0, const float *rpz0, const float *rpxEnd,
                      const float *rpyEnd, const float *rpzEnd)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double dt = g->dt, hx = g->dx, hy = g->dy, hz = g->dz, m = g->mu0, e = g->eps0;
    const size_t sy = nx, sz = (size_t)nx * ny;
    for (int k = 1; k < nz - 1; k++)
        for (int i = 0; i < nx - 1; i++) {
            size_t o = IDX(i, 0, k);
            exT[o] = mur64(dt, hy, m, e, RPY(rpy0, i, 0, k), RPY(rpy0, i, 1, k), exT[o + sy], exS[o + sy], exS[o]);
            o = IDX(i, ny - 1, k);
            exT[o] = mur64(dt, hy, m, e, RPY(rpyEnd, i, 1, k), RPY(rpyEnd, i, 0, k), exT[o - sy], exS[o - sy], exS[o]);
        }
    for (int j = 1; j < ny - 1; j++)
        for (int i = 0; i < nx - 1; i++) {
            size_t o = IDX(i, j, 0);
            exT[o] = mur64(dt, hz, m, e, RPZ(rpz0, i, j, 0), RPZ(rpz0, i, j, 1), exT[o + sz], exS[o + sz], exS[o]);
            o = IDX(i, j, nz - 1);
            exT[o] = mur64(dt, hz, m, e, RPZ(rpzEnd, i, j, 1), RPZ(rpzEnd, i, j, 0), exT[o - sz], exS[o - sz], exS[o]);
        }
    for (int k = 1; k < nz - 1; k++)
        for (int j = 0; j < ny - 1; j++) {
            size_t o = IDX(0, j, k);
            eyT[o] = mur64(dt, hx, m, e, RPX(rpx0, 0, j, k), RPX(rpx0, 1, j, k), eyT[o + 1], eyS[o + 1], eyS[o]);
            o = IDX(nx - 1, j, k);
            eyT[o] = mur64(dt, hx, m, e, RPX(rpxEnd, 1, j, k), RPX(rpxEnd, 0, j, k), eyS[o - 1], eyS[o - 1], eyS[o]);
        }
    for (int j = 0; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            size_t o = IDX(i, j, 0);
            eyT[o] = mur64(dt, hz, m, e, RPZ(rpz0, i, j, 0), RPZ(rpz0, i, j, 1), eyT[o + sz], eyS[o + sz], eyS[o]);
            o = IDX(i, j, nz - 1);
            eyT[o] = mur64(dt, hz, m, e, RPZ(rpzEnd, i, j, 1), RPZ(rpzEnd, i, j, 0), eyT[o - sz], eyS[o - sz], eyS[o]);
        }
    for (int k = 0; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++) {
            size_t o = IDX(0, j, k);
            ezT[o] = mur64(dt, hx, m, e, RPX(rpx0, 0, j, k), RPX(rpx0, 1, j, k), ezT[o + 1], ezS[o + 1], ezS[o]);
            o = IDX(nx - 1, j, k);
            ezT[o] = mur64(dt, hx, m, e, RPX(rpxEnd, 1, j, k), RPX(rpxEnd, 0, j, k), ezT[o - 1], ezS[o - 1], ezS[o]);
        }
    for (int k = 0; k < nz - 1; k++)
        for (int i = 1; i < nx - 1; i++) {
            size_t o = IDX(i, 0, k);
            ezT[o] = mur64(dt, hy, m, e, RPY(rpy0, i, 0, k), RPY(rpy0, i, 1, k), ezT[o + sy], ezS[o + sy], ezS[o]);
            o = IDX(i, ny - 1, k);
            ezT[o] = mur64(dt, hy, m, e, RPY(rpyEnd, i, 1, k), RPY(rpyEnd, i, 0, k), ezT[o - sy], ezS[o - sy], ezS[o]);
        }
}

/* ------------------------------------------------------------------ */
/* Whole-run simulation object: the step sequence of fdtd.cu:115-530  */
/* ------------------------------------------------------------------ */

/* Field ids (the column order of the reference's probe writer, fdtd.cu:1223):
 * 0..2 Dx Dy Dz, 3..5 Hx Hy Hz, 6..8 Ex Ey Ez. */
typedef struct {
    or_grid g;
    or_grid64 g64;
    int prec;             /* 32 or 64 */
    int debye_mode;
    int mur_on;           /* 1: apply the Mur kernel each step (the reference always does) */
    int n;                /* next step index */
    size_t N;
    void *E[3][3], *D[3][3], *H[3];   /* [slot][component] */
    float *sigma, *epsI, *epsS, *tau;
    float *rp[6];
    int nsrc; int *src;
    int jz_len; float *jz; int jz_mode;   /* jz_mode 0: jz[3*floor(n/3)] (fdtd.cu:172); 1: jz[n] */
} or_sim;

or_sim *or_sim_create(int nx, int ny, int nz, float dt, float dx, float dy, float dz, float mu0, float eps0, int prec)
{
    or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
    s->g.nx = nx; s->g.ny = ny; s->g.nz = nz;
    s->g.dt = dt; s->g.dx = dx; s->g.dy = dy; s->g.dz = dz; s->g.mu0 = mu0; s->g.eps0 = eps0;
    grid64(&s->g, &s->g64);
    s->prec = prec;
    s->debye_mode = OR_DEBYE_FIXED;
    s->mur_on = 1;
    s->N = (size_t)nx * ny * nz;
    size_t es = (prec == 64) ? sizeof(double) : sizeof(float);
    for (int a = 0; a < 3; a++) {
        s->H[a] = calloc(s->N, es);
        for (int b = 0; b < 3; b++) { s->E[a][b] = calloc(s->N, es); s->D[a][b] = calloc(s->N, es); }
    }
    s->sigma = (float *)calloc(s->N, sizeof(float));
    s->epsI = (float *)calloc(s->N, sizeof(float));
    s->epsS = (float *)calloc(s->N, sizeof(float));
    s->tau = (float *)calloc(s->N, sizeof(float));
    size_t rps[6] = {(size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2,
                     (size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2};
    for (int q = 0; q < 6; q++) {
        s->rp[q] = (float *)malloc(rps[q] * sizeof(float));
        for (size_t t = 0; t < rps[q]; t++) s->rp[q][t] = 1.0f;
    }
    for (size_t t = 0; t < s->N; t++) { s->epsI[t] = 1.0f; s->epsS[t] = 1.0f; }
    return s;
}

void or_sim_destroy(or_sim *s)
{
    if (!s) return;
    for (int a = 0; a < 3; a++) {
        free(s->H[a]);
        for (int b = 0; b < 3; b++) { free(s->E[a][b]); free(s->D[a][b]); }
    }
    free(s->sigma); free(s->epsI); free(s->epsS); free(s->tau);
    for (int q = 0; q < 6; q++) free(s->rp[q]);
    free(s->src); free(s->jz);
    free(s);
}

void or_sim_set_modes(or_sim *s, int debye_mode, int mur_on) { s->debye_mode = debye_mode; s->mur_on = mur_on; }

void or_sim_set_materials(or_sim *s, const float *sigma, const float *epsS, const float *epsI, const float *tau)
{
    memcpy(s->sigma, sigma, s->N * sizeof(float));
    memcpy(s->epsS, epsS, s->N * sizeof(float));
    memcpy(s->epsI, epsI, s->N * sizeof(float));
    memcpy(s->tau, tau, s->N * sizeof(float));
}

/* order: rpx0, rpy0, rpz0, rpxEnd, rpyEnd, rpzEnd */
void or_sim_set_rp(or_sim *s, int which, const float *rp)
{
    const int nx = s->g.nx, ny = s->g.ny, nz = s->g.nz;
    size_t rps[6] = {(size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2,
                     (size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2};
    memcpy(s->rp[which], rp, rps[which] * sizeof(float));
}

void or_sim_set_sources(or_sim *s, int nsrc, const int *xyz, int jz_len, const float *jz, int jz_mode)
{
    free(s->src); free(s->jz);
    s->nsrc = nsrc;
    s->src = (int *)malloc(sizeof(int) * 3 * (nsrc > 0 ? nsrc : 1));
    if (nsrc > 0) memcpy(s->src, xyz, sizeof(int) * 3 * nsrc);
    s->jz_len = jz_len;
    s->jz = (float *)malloc(sizeof(float) * (jz_len > 0 ? jz_len : 1));
    if (jz_len > 0) memcpy(s->jz, jz, sizeof(float) * jz_len);
    s->jz_mode = jz_mode;
}

static float jz_at(const or_sim *s, int n)
{
    int idx = (s->jz_mode == 0) ? (n / 3) * 3 : n;
    return (idx >= 0 && idx < s->jz_len) ? s->jz[idx] : 0.0f;
}

/* level 0 = newest (E^n, D^n after n steps), 1 = previous, 2 = oldest slot */
static int slot_of(const or_sim *s, int level) { return ((s->n + 2 - level) % 3 + 3) % 3; }

void or_sim_step(or_sim *s)
{
    int n = s->n;
    int t = n % 3, s1 = (n + 2) % 3, s2 = (n + 1) % 3;
    if (s->prec == 32) {
        float *Et[3], *E0[3], *E1[3], *Dt[3], *D0[3], *D1[3], *H[3];
        for (int q = 0; q < 3; q++) {
            Et[q] = (float *)s->E[t][q]; E0[q] = (float *)s->E[s1][q]; E1[q] = (float *)s->E[s2][q];
            Dt[q] = (float *)s->D[t][q]; D0[q] = (float *)s->D[s1][q]; D1[q] = (float *)s->D[s2][q];
            H[q] = (float *)s->H[q];
        }
        or_update_h_f32(&s->g, H[0], H[1], H[2], E0[0], E0[1], E0[2]);
        or_update_d_f32(&s->g, Dt[0], Dt[1], Dt[2], D0[0], D0[1], D0[2], H[0], H[1], H[2]);
        or_update_sources_f32(&s->g, Dt[2], D0[2], H[0], H[1], s->nsrc, s->src, jz_at(s, n));
        or_update_e_f32(&s->g, s->debye_mode, Et[0], Et[1], Et[2], E0[0], E0[1], E0[2], E1[0], E1[1], E1[2],
                        Dt[0], Dt[1], Dt[2], D0[0], D0[1], D0[2], D1[0], D1[1], D1[2],
                        s->sigma, s->epsI, s->epsS, s->tau);
        if (s->mur_on)
            or_update_mur_f32(&s->g, Et[0], Et[1], Et[2], E0[0], E0[1], E0[2],
                              s->rp[0], s->rp[1], s->rp[2], s->rp[3], s->rp[4], s->rp[5]);
    } else {
        double *Et[3], *E0[3], *E1[3], *Dt[3], *D0[3], *D1[3], *H[3];
        for (int q = 0; q < 3; q++) {
            Et[q] = (double *)s->E[t][q]; E0[q] = (double *)s->E[s1][q]; E1[q] = (double *)s->E[s2][q];
            Dt[q] = (double *)s->D[t][q]; D0[q] = (double *)s->D[s1][q]; D1[q] = (double *)s->D[s2][q];
            H[q] = (double *)s->H[q];
        }
        upd_h64(&s->g64, H[0], H[1], H[2], E0[0], E0[1], E0[2]);
        upd_d64(&s->g64, Dt[0], Dt[1], Dt[2], D0[0], D0[1], D0[2], H[0], H[1], H[2]);
        upd_src64(&s->g64, Dt[2], D0[2], H[0], H[1], s->nsrc, s->src, (double)jz_at(s, n));
        upd_e64(&s->g64, s->debye_mode, Et, E0, E1, Dt, D0, D1, s->sigma, s->epsI, s->epsS, s->tau);
        if (s->mur_on)
            upd_mur64(&s->g64, Et[0], Et[1], Et[2], E0[0], E0[1], E0[2],
                      s->rp[0], s->rp[1], s->rp[2], s->rp[3], s->rp[4], s->rp[5]);
    }
    s->n = n + 1;
}

static void *field_ptr(or_sim *s, int field, int level)
{
    int q = field % 3;
    if (field >= 3 && field < 6) return s->H[q];
    int sl = slot_of(s, level);
    return field < 3 ? s->D[sl][q] : s->E[sl][q];
}

static double read_cell(or_sim *s, int field, int level, size_t o)
{
    void *p = field_ptr(s, field, level);
    return s->prec == 64 ? ((double *)p)[o] : (double)((float *)p)[o];
}

/* Runs `steps` steps; after each, records probes (field, i, j, k) as double
 * into out[step * nprobe + p]. */
void or_sim_run(or_sim *s, int steps, int nprobe, const int *probes, double *out)
{
    const int nx = s->g.nx, ny = s->g.ny;
    for (int t = 0; t < steps; t++) {
        or_sim_step(s);
        for (int p = 0; p < nprobe; p++) {
            const int *q = probes + 4 * p;
            out[(size_t)t * nprobe + p] = read_cell(s, q[0], 0, IDX(q[1], q[2], q[3]));
        }
    }
}

int or_sim_steps_done(const or_sim *s) { return s->n; }

void or_sim_get(or_sim *s, int field, int level, double *out)
{
    void *p = field_ptr(s, field, level);
    if (s->prec == 64) memcpy(out, p, s->N * sizeof(double));
    else for (size_t t = 0; t < s->N; t++) out[t] = ((float *)p)[t];
}

void or_sim_get_f32(or_sim *s, int field, int level, float *out)
{
    void *p = field_ptr(s, field, level);
    if (s->prec == 32) memcpy(out, p, s->N * sizeof(float));
    else for (size_t t = 0; t < s->N; t++) out[t] = (float)((double *)p)[t];
}

void or_sim_set(or_sim *s, int field, int level, const double *in)
{
    void *p = field_ptr(s, field, level);
    if (s->prec == 64) memcpy(p, in, s->N * sizeof(double));
    else for (size_t t = 0; t < s->N; t++) ((float *)p)[t] = (float)in[t];
}

void or_sim_set_f32(or_sim *s, int field, int level, const float *in)
{
    void *p = field_ptr(s, field, level);
    if (s->prec == 32) memcpy(p, in, s->N * sizeof(float));
    else for (size_t t = 0; t < s->N; t++) ((double *)p)[t] = in[t];
}

/* Sum over all cells of E.D/2 + mu0 H.H/2 with the newest level, in double
 * (used by the absorbing-boundary energy checks). */
double or_sim_energy(or_sim *s)
{
    double acc = 0.0;
    for (int q = 0; q < 3; q++)
        for (size_t t = 0; t < s->N; t++) {
            double e = read_cell(s, 6 + q, 0, t), d = read_cell(s, q, 0, t), h = read_cell(s, 3 + q, 0, t);
            acc += 0.5 * e * d + 0.5 * (double)s->g.mu0 * h * h;
        }
    return acc;
}
