// Host-side services of the FDTD path (fdtd.h): parameter file, tissue
// table and phantom slices, Mur permittivity planes, source waveform,
// host/device allocation, initial upload and the text result writers.
//
// Each function keeps the name, arguments and observable behaviour of the
// reference function cited above it; the arithmetic of the derived constants,
// the waveform and the rp planes uses the same operand types as the reference
// so the values agree to the bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fdtd.h"
#include "fdtd_amd.h"
#include "fdtd_internal.h"

#define HCHECK(call)                                                            \
    do {                                                                        \
        const hipError_t error_ = (call);                                       \
        if (error_ != hipSuccess) {                                             \
            printf("Error: %s: %d\n", __FILE__, __LINE__);                      \
            printf("Code %d: %s\n", (int)error_, hipGetErrorString(error_));    \
            exit(EXIT_FAILURE);                                                 \
        }                                                                       \
    } while (0)

namespace {

[[noreturn]] void die(const char *msg, const char *arg)
{
    printf(msg, arg);
    exit(EXIT_FAILURE);
}

// Reads one whitespace-delimited token; returns false at end of file.
bool token(FILE *f, std::string &out)
{
    out.clear();
    int ch;
    while ((ch = fgetc(f)) != EOF && isspace(ch)) {}
    if (ch == EOF) return false;
    do { out.push_back((char)ch); } while ((ch = fgetc(f)) != EOF && !isspace(ch));
    if (ch != EOF) ungetc(ch, f);
    return true;
}

// Skips the rest of the current line (the reference's fgets of an unused key).
void skip_line(FILE *f)
{
    int ch;
    while ((ch = fgetc(f)) != EOF && ch != '\n') {}
}

void skip_ws(FILE *f)
{
    int ch;
    while ((ch = fgetc(f)) != EOF && isspace(ch)) {}
    if (ch != EOF) ungetc(ch, f);
}

bool read_key_values(FILE *f, int n, std::string *vals)
{
    std::string key;
    if (!token(f, key)) return false;
    for (int q = 0; q < n; q++)
        if (!token(f, vals[q])) return false;
    skip_ws(f);
    return true;
}

}  // namespace

// fdtd.cu:547-637.  The reference reads 20 positional lines (12 of them
// skipped) and stops after ftau_d; lines after it are extension keys here
// (the reference ignores them): "courant_factor <s>" sets timeskip (the
// reference hard-codes 1.0 at fdtd.cu:627) and "cell_size <m>" sets
// dx = dy = dz instead of lambda / elements_per_wavelength.
extern "C" FdtdParams *initParamsWithPath(const char *filePath)
{
    FILE *f = fopen(filePath, "r");
    if (!f) die("Couldn't open file %s\n", filePath);
    FdtdParams *p = (FdtdParams *)calloc(1, sizeof(FdtdParams));
    p->inputPath = (char *)calloc(1024, 1);
    p->outputPath = (char *)calloc(1024, 1);
    std::string v[3];
    bool ok = true;
    ok = ok && read_key_values(f, 3, v);
    if (ok) { p->nx = atoi(v[0].c_str()); p->ny = atoi(v[1].c_str()); p->nz = atoi(v[2].c_str()); }
    ok = ok && read_key_values(f, 1, v);
    if (ok) p->iterationsCount = ((atoi(v[0].c_str()) - 1) / 3 + 1) * 3;  // multiple of 3 (fdtd.cu:563)
    if (ok) skip_line(f);                                                 // nf
    ok = ok && read_key_values(f, 1, v);
    if (ok) snprintf(p->inputPath, 1024, "%s", v[0].c_str());
    if (ok) skip_line(f);                                                 // env_file_prefix
    ok = ok && read_key_values(f, 1, v);
    if (ok) snprintf(p->outputPath, 1024, "%s", v[0].c_str());
    for (int u = 0; ok && u < 9; u++) skip_line(f);                       // output_format .. source_type
    ok = ok && read_key_values(f, 1, v);
    if (ok) p->elementsPerWave = atoi(v[0].c_str());
    ok = ok && read_key_values(f, 1, v);
    if (ok) p->waveFrequency = strtof(v[0].c_str(), nullptr);
    ok = ok && read_key_values(f, 1, v);
    if (ok) p->pulseWidth = strtof(v[0].c_str(), nullptr);
    ok = ok && read_key_values(f, 1, v);
    if (ok) p->pulseModulationFrequency = strtof(v[0].c_str(), nullptr);
    ok = ok && read_key_values(f, 1, v);
    if (ok) p->sourcesCount = atoi(v[0].c_str());
    if (ok && (p->sourcesCount < 0 || p->sourcesCount > (1 << 20))) ok = false;
    if (ok) {
        p->sources = (int *)calloc((size_t)3 * (p->sourcesCount > 0 ? p->sourcesCount : 1), sizeof(int));
        for (int s = 0; ok && s < p->sourcesCount; s++) {
            ok = read_key_values(f, 3, v);
            for (int q = 0; ok && q < 3; q++) p->sources[3 * s + q] = atoi(v[q].c_str()) - 1;  // 1-based in the file
        }
    }
    if (ok) skip_line(f);  // pulse_type
    float *defaults[4] = {&p->defaultSigma, &p->defaultEpsS, &p->defaultEpsI, &p->defaultTauD};
    for (int q = 0; ok && q < 4; q++) {
        ok = read_key_values(f, 1, v);
        if (ok) *defaults[q] = strtof(v[0].c_str(), nullptr);
    }
    if (!ok) {
        fclose(f);
        die("Malformed parameter file %s\n", filePath);
    }
    float timeskip = 1.0f, cell = 0.0f;
    std::string key;
    while (token(f, key)) {
        std::string val;
        if (!key.empty() && key.back() == ':') key.pop_back();
        if (key == "courant_factor" && token(f, val)) timeskip = strtof(val.c_str(), nullptr);
        else if (key == "cell_size" && token(f, val)) cell = strtof(val.c_str(), nullptr);
        skip_line(f);
    }
    fclose(f);

    p->pi = acos(-1.0);
    p->c = 3.0 * pow(10.0, 8.0);
    p->timeskip = timeskip;
    p->lambda = p->c / p->waveFrequency;
    p->dx = cell > 0.0f ? cell : p->lambda / p->elementsPerWave;
    p->dy = p->dx;
    p->dz = p->dx;
    p->dt = 1.0 * p->timeskip / (p->c * sqrt(1.0 / pow(p->dx, 2.0) + 1.0 / pow(p->dy, 2.0) + 1.0 / pow(p->dz, 2.0)));
    p->mu0 = 4.0 * p->pi * pow(10.0, -7.0);
    p->eps0 = 1.0 / (p->mu0 * p->c * p->c);
    return p;
}

extern "C" void deallocParams(FdtdParams *params)
{
    if (!params) return;
    free(params->inputPath);
    free(params->outputPath);
    free(params->sources);
    free(params->jz);
    free(params);
}

// fdtd.cu:648-674
extern "C" void printParams(FdtdParams *params)
{
    printf("Field size:                 %04dx%04dx%04d\n", params->nx, params->ny, params->nz);
    printf("Iterations count:           %d\n", params->iterationsCount);
    printf("Input path:                 %s\n", params->inputPath);
    printf("Output path:                %s\n", params->outputPath);
    printf("Elements per wavelength:    %d\n", params->elementsPerWave);
    printf("Wave frequency:             %9.3E\n", params->waveFrequency);
    printf("Pulse width:                %9.3E\n", params->pulseWidth);
    printf("Pulse modulation frequency: %9.3E\n", params->pulseModulationFrequency);
    printf("Sources count:              %d\n", params->sourcesCount);
    for (int i = 0; i < params->sourcesCount; i++)
        printf("Source position:            %04dx%04dx%04d\n", params->sources[i * 3 + 0] + 1,
               params->sources[i * 3 + 1] + 1, params->sources[i * 3 + 2] + 1);
    printf("Default sigma:              %9.3E\n", params->defaultSigma);
    printf("Default eps_s:              %9.3E\n", params->defaultEpsS);
    printf("Default eps_i:              %9.3E\n", params->defaultEpsI);
    printf("Default tau_d:              %9.3E\n", params->defaultTauD);
    printf("\n");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) {
        printf("Running on %s\n", prop.name);
        printf("Architecture: %s\n", prop.gcnArchName);
    }
}

// fdtd.cu:677-725
extern "C" FdtdField *initFieldWithParams(FdtdParams *params)
{
    const size_t n = (size_t)params->nx * params->ny * params->nz;
    FdtdField *field = (FdtdField *)calloc(1, sizeof(FdtdField));
    if (!field) die("Couldn't allocate field%s\n", "");
    float **pinned[9] = {&field->hx, &field->hy, &field->hz, &field->dx0, &field->dy0, &field->dz0,
                         &field->ex0, &field->ey0, &field->ez0};
    for (auto pp : pinned) HCHECK(hipHostMalloc((void **)pp, n * sizeof(float), hipHostMallocDefault));
    field->sigma = (float *)malloc(n * sizeof(float));
    field->epsS = (float *)malloc(n * sizeof(float));
    field->epsI = (float *)malloc(n * sizeof(float));
    field->tauD = (float *)malloc(n * sizeof(float));
    for (size_t i = 0; i < n; i++) {
        field->sigma[i] = params->defaultSigma;
        field->epsS[i] = params->defaultEpsS;
        field->epsI[i] = params->defaultEpsI;
        field->tauD[i] = params->defaultTauD;
    }
    const size_t nx = params->nx, ny = params->ny, nz = params->nz;
    field->rpx0 = (float *)malloc(2 * ny * nz * sizeof(float));
    field->rpy0 = (float *)malloc(nx * 2 * nz * sizeof(float));
    field->rpz0 = (float *)malloc(nx * ny * 2 * sizeof(float));
    field->rpxEnd = (float *)malloc(2 * ny * nz * sizeof(float));
    field->rpyEnd = (float *)malloc(nx * 2 * nz * sizeof(float));
    field->rpzEnd = (float *)malloc(nx * ny * 2 * sizeof(float));
    return field;
}

// fdtd.cu:728-761
extern "C" void deallocField(FdtdField *field)
{
    if (!field) return;
    float *pinned[9] = {field->hx, field->hy, field->hz, field->dx0, field->dy0, field->dz0,
                        field->ex0, field->ey0, field->ez0};
    for (auto p : pinned) HCHECK(hipHostFree(p));
    free(field->sigma); free(field->epsS); free(field->epsI); free(field->tauD);
    free(field->rpx0); free(field->rpy0); free(field->rpz0);
    free(field->rpxEnd); free(field->rpyEnd); free(field->rpzEnd);
    free(field);
}

// fdtd.cu:764-822
extern "C" FdtdField *initDeviceFieldWithParams(FdtdParams *params)
{
    const size_t n = (size_t)params->nx * params->ny * params->nz;
    const size_t nx = params->nx, ny = params->ny, nz = params->nz;
    FdtdField *field = (FdtdField *)calloc(1, sizeof(FdtdField));
    float **full[25] = {&field->ex0, &field->ey0, &field->ez0, &field->ex1, &field->ey1, &field->ez1,
                        &field->ex2, &field->ey2, &field->ez2, &field->hx, &field->hy, &field->hz,
                        &field->dx0, &field->dy0, &field->dz0, &field->dx1, &field->dy1, &field->dz1,
                        &field->dx2, &field->dy2, &field->dz2, &field->epsI, &field->epsS, &field->tauD,
                        &field->sigma};
    for (auto pp : full) HCHECK(hipMalloc((void **)pp, n * sizeof(float)));
    HCHECK(hipMalloc((void **)&field->rpx0, 2 * ny * nz * sizeof(float)));
    HCHECK(hipMalloc((void **)&field->rpy0, nx * 2 * nz * sizeof(float)));
    HCHECK(hipMalloc((void **)&field->rpz0, nx * ny * 2 * sizeof(float)));
    HCHECK(hipMalloc((void **)&field->rpxEnd, 2 * ny * nz * sizeof(float)));
    HCHECK(hipMalloc((void **)&field->rpyEnd, nx * 2 * nz * sizeof(float)));
    HCHECK(hipMalloc((void **)&field->rpzEnd, nx * ny * 2 * sizeof(float)));
    size_t memFree = 0, memTotal = 0;
    if (hipMemGetInfo(&memFree, &memTotal) == hipSuccess) {
        printf("Memory available: %.2f MB\n", (float)memTotal / (1024.0 * 1024.0));
        printf("Memory used: %.2f MB\n\n", (float)(memTotal - memFree) / (1024.0 * 1024.0));
    }
    return field;
}

// fdtd.cu:825-872
extern "C" void deallocDeviceField(FdtdField *field)
{
    if (!field) return;
    float *all[31] = {field->ex0, field->ey0, field->ez0, field->ex1, field->ey1, field->ez1, field->ex2,
                      field->ey2, field->ez2, field->hx, field->hy, field->hz, field->dx0, field->dy0,
                      field->dz0, field->dx1, field->dy1, field->dz1, field->dx2, field->dy2, field->dz2,
                      field->epsI, field->epsS, field->tauD, field->sigma, field->rpx0, field->rpy0,
                      field->rpz0, field->rpxEnd, field->rpyEnd, field->rpzEnd};
    for (auto p : all) HCHECK(hipFree(p));
    free(field);
}

namespace {

// Re(eps(omega)) of a Debye medium in the two forms of fdtd.cu:959-1019.
// x planes: real c1 and c2 (no I), so the conductivity term and the
// (1 + omega tau) denominator are real.  y/z planes: c1 = j omega tau and
// c2 = j sigma / (omega eps0), whose real part carries no conductivity.
// float-complex operands promoted to double-complex, as C evaluates them.
float rp_value(bool imaginary, float f, float pi, float eps0, float sg, float es, float ei, float tau)
{
    const double w = 2.0 * pi * f * tau;
    const double s = sg / (2.0 * pi * f * eps0);
    const std::complex<float> c1 = imaginary ? std::complex<float>(0.0f, (float)w) : std::complex<float>((float)w, 0.0f);
    const std::complex<float> c2 = imaginary ? std::complex<float>(0.0f, (float)s) : std::complex<float>((float)s, 0.0f);
    const std::complex<double> den(1.0 + (double)c1.real(), (double)c1.imag());
    const std::complex<double> q = std::complex<double>((double)(es - ei), 0.0) / den;
    const std::complex<double> r = std::complex<double>((double)ei, 0.0) + q - std::complex<double>(c2);
    return (float)r.real();
}

}  // namespace

// fdtd.cu:942-1024: two cells nearest each face, from the field's material
// arrays at the time of the call.
extern "C" void setupMurBoundary(FdtdParams *params, FdtdField *field)
{
    const int nx = params->nx, ny = params->ny, nz = params->nz;
    const float f = params->waveFrequency, pi = params->pi, e0 = params->eps0;
    auto at = [&](int i, int j, int k) { return ((size_t)k * ny + j) * nx + i; };
    auto rp = [&](bool im, size_t o) {
        return rp_value(im, f, pi, e0, field->sigma[o], field->epsS[o], field->epsI[o], field->tauD[o]);
    };
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int s = 0; s < 2; s++) {
                field->rpx0[((size_t)k * ny + j) * 2 + s] = rp(false, at(s, j, k));
                field->rpxEnd[((size_t)k * ny + j) * 2 + s] = rp(false, at(nx - 2 + s, j, k));
            }
    for (int k = 0; k < nz; k++)
        for (int s = 0; s < 2; s++)
            for (int i = 0; i < nx; i++) {
                field->rpy0[((size_t)k * 2 + s) * nx + i] = rp(true, at(i, s, k));
                field->rpyEnd[((size_t)k * 2 + s) * nx + i] = rp(true, at(i, ny - 2 + s, k));
            }
    for (int s = 0; s < 2; s++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                field->rpz0[((size_t)s * ny + j) * nx + i] = rp(true, at(i, j, s));
                field->rpzEnd[((size_t)s * ny + j) * nx + i] = rp(true, at(i, j, nz - 2 + s));
            }
}

// fdtd.cu:875-939.  Tissue rows "code name sigma eps_s eps_inf tau" are read
// until a code >= 94 (stored, then reading stops) or 94 rows; then nz slices
// <materialsPath>/v1_%05d.pgm, each a 6-token header followed by nx*ny codes
// read in order whatever the image width (the reference's shear).  The
// parsers are those of fdtd_load_tissue_table / fdtd_load_phantom_ids.
extern "C" void loadMaterials(FdtdParams *params, FdtdField *field, const char *specsFilePath,
                              const char *materialsPath)
{
    std::vector<float> specs(95 * 4, 0.0f);
    std::string err;
    if (phantom_read_table(specsFilePath, specs.data(), err) < 0) die("%s\n", err.c_str());
    const int nx = params->nx, ny = params->ny, nz = params->nz;
    const size_t nxy = (size_t)nx * ny;
    std::vector<unsigned char> ids(nxy);
    std::vector<char> buf;
    for (int iz = 0; iz < nz; iz++) {
        char name[1100];
        snprintf(name, sizeof name, "%s/v1_%05d.pgm", materialsPath, iz + 1);
        if (phantom_read_slice(name, nx, ny, FDTD_PGM_REFERENCE, ids.data(), buf, err) < 0) die("%s\n", err.c_str());
        float *sg = field->sigma + iz * nxy, *es = field->epsS + iz * nxy, *ei = field->epsI + iz * nxy,
              *tau = field->tauD + iz * nxy;
        for (size_t t = 0; t < nxy; t++) {
            const int code = ids[t];
            sg[t] = specs[code * 4 + 0];
            es[t] = specs[code * 4 + 1];
            ei[t] = specs[code * 4 + 2];
            tau[t] = specs[code * 4 + 3];
        }
    }
}

// fdtd.cu:1027-1082: a Gaussian times a cosine carrier, sampled every
// `fine` sub-steps from the first sample of magnitude >= 1e-9, scaled by
// 1e-15/dt/3, then centrally differenced and scaled by dt*dz/(dx*dy*dz).
extern "C" void setupSources(FdtdParams *params)
{
    const int off = 1 << 16;
    params->jzCount = off;
    free(params->jz);
    params->jz = (float *)calloc(off, sizeof(float));
    std::vector<float> g((size_t)off * 2, 0.0f), d((size_t)off * 2, 0.0f);
    const float pw = params->pulseWidth, f = params->waveFrequency, pmf = params->pulseModulationFrequency;
    const float dt = params->dt;
    const int fine = (1 << 13) * pw * f * dt;
    if (fine <= 0) die("Source waveform: sampling step is zero%s\n", "");
    const int temp = 1.0 / (pw * f) / (dt / fine) / 2.0;
    if (temp + 1 >= off) die("Source waveform: pulse too long%s\n", "");
    for (int i = -temp - 1; i <= temp + 1; i++) {
        const float v1 = ((float)i / (((float)temp + 1.0) / 4.0));
        const float v2 = exp(-pow(v1, 2.0));
        const float v3 = cos(2.0 * acos(-1.0) * pmf * f * i * (dt / fine));
        g[i + off] = v2 * v3;
    }
    int istart = 0;
    bool found = false;
    for (int i = -(1 << 12); i < (1 << 12) && !found; i++)
        if (std::fabs((double)g[i + off]) >= 1e-9 && i % fine == 0) { istart = i; found = true; }
    if (!found) die("Source waveform: no start sample%s\n", "");
    float *jz = params->jz;
    int n = 0;
    for (int i = istart; i <= temp + 1; i += fine) {
        if (n >= off) die("Source waveform: too many samples%s\n", "");
        jz[n++] = g[i + off] * 1e-15 / dt / 3.0;
    }
    for (int i = 2; i <= (1 << 14); i++)
        d[i - 1 + off] = (((jz[i] - jz[i - 1]) / dt) + ((jz[i - 1] - jz[i - 2]) / dt)) / 2.0 *
                         (params->dt * params->dz) / (params->dx * params->dy * params->dz);
    for (int i = 0; i < (1 << 14); i++) jz[i] = d[i + off + 1];
}

// fdtd.cu:1085-1133
extern "C" void copyDataToDevice(FdtdParams *params, FdtdField *field, FdtdField *deviceField)
{
    const size_t n = (size_t)params->nx * params->ny * params->nz;
    const size_t bytes = n * sizeof(float);
    float *zero[21] = {deviceField->hx, deviceField->hy, deviceField->hz, deviceField->dx0, deviceField->dy0,
                       deviceField->dz0, deviceField->dx1, deviceField->dy1, deviceField->dz1, deviceField->dx2,
                       deviceField->dy2, deviceField->dz2, deviceField->ex0, deviceField->ey0, deviceField->ez0,
                       deviceField->ex1, deviceField->ey1, deviceField->ez1, deviceField->ex2, deviceField->ey2,
                       deviceField->ez2};
    for (auto p : zero) HCHECK(hipMemset(p, 0, bytes));
    HCHECK(hipMemcpy(deviceField->epsI, field->epsI, bytes, hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(deviceField->epsS, field->epsS, bytes, hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(deviceField->tauD, field->tauD, bytes, hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(deviceField->sigma, field->sigma, bytes, hipMemcpyHostToDevice));
    const size_t nx = params->nx, ny = params->ny, nz = params->nz;
    HCHECK(hipMemcpy(deviceField->rpx0, field->rpx0, 2 * ny * nz * sizeof(float), hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(deviceField->rpy0, field->rpy0, nx * 2 * nz * sizeof(float), hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(deviceField->rpz0, field->rpz0, nx * ny * 2 * sizeof(float), hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(deviceField->rpxEnd, field->rpxEnd, 2 * ny * nz * sizeof(float), hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(deviceField->rpyEnd, field->rpyEnd, nx * 2 * nz * sizeof(float), hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(deviceField->rpzEnd, field->rpzEnd, nx * ny * 2 * sizeof(float), hipMemcpyHostToDevice));
}

// fdtd.cu:1136-1160: waits for the stream's D2H into the shared pinned
// buffers, snapshots them into fresh buffers and releases copyMutex.
extern "C" void *copyResultsWithParams(void *arg)
{
    CopyParams *cp = (CopyParams *)arg;
    const size_t bytes = (size_t)cp->params->nx * cp->params->ny * cp->params->nz * sizeof(float);
    cp->xBuffer = (float *)malloc(bytes);
    cp->yBuffer = (float *)malloc(bytes);
    cp->zBuffer = (float *)malloc(bytes);
    if (!cp->xBuffer || !cp->yBuffer || !cp->zBuffer) die("Couldn't allocate result buffers%s\n", "");
    HCHECK(hipStreamSynchronize(cp->stream));
    memcpy(cp->xBuffer, cp->xSource, bytes);
    memcpy(cp->yBuffer, cp->ySource, bytes);
    memcpy(cp->zBuffer, cp->zSource, bytes);
    pthread_mutex_unlock(cp->copyMutex);
    return nullptr;
}

// fdtd.cu:1163-1194 (the reference's throttle counter is the caller's).
extern "C" void *writeResultsWithParams(void *arg)
{
    ResultsParams *rp = (ResultsParams *)arg;
    pthread_mutex_lock(rp->hParams->copyMutex);
    pthread_mutex_lock(rp->dParams->copyMutex);
    pthread_mutex_lock(rp->eParams->copyMutex);
    writeResults3d(rp->params, rp->hParams->xBuffer, rp->hParams->yBuffer, rp->hParams->zBuffer,
                   rp->dParams->xBuffer, rp->dParams->yBuffer, rp->dParams->zBuffer, rp->eParams->xBuffer,
                   rp->eParams->yBuffer, rp->eParams->zBuffer, rp->currentIteration);
    pthread_mutex_unlock(rp->hParams->copyMutex);
    pthread_mutex_unlock(rp->dParams->copyMutex);
    pthread_mutex_unlock(rp->eParams->copyMutex);
    free(rp->hParams->copyMutex);
    free(rp->dParams->copyMutex);
    free(rp->eParams->copyMutex);
    free(rp->hParams);
    free(rp->dParams);
    free(rp->eParams);
    free(rp);
    return nullptr;
}

namespace {

FILE *open_out(const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f) die("Couldn't open file %s\n", path);
    return f;
}

}  // namespace

// fdtd.cu:1197-1285: for each source, the x, y and z lines through it;
// rows " %3d %3d %3d" (1-based) then Dx Dy Dz Hx Hy Hz Ex Ey Ez as " % 9.3E".
// Frees the nine buffers, as the reference does.
extern "C" void writeResults(FdtdParams *params, float *hxSource, float *hySource, float *hzSource,
                             float *dxSource, float *dySource, float *dzSource, float *exSource,
                             float *eySource, float *ezSource, int currentIteration)
{
    const int nx = params->nx, ny = params->ny, nz = params->nz;
    const float *F[9] = {dxSource, dySource, dzSource, hxSource, hySource, hzSource, exSource, eySource, ezSource};
    char path[1100];
    for (int axis = 0; axis < 3; axis++) {
        snprintf(path, sizeof path, "%s/E_field_%c_%05d.out", params->outputPath, "xyz"[axis], currentIteration + 1);
        FILE *f = open_out(path);
        for (int s = 0; s < params->sourcesCount; s++) {
            int c[3] = {params->sources[3 * s], params->sources[3 * s + 1], params->sources[3 * s + 2]};
            const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
            for (int t = 0; t < len; t++) {
                c[axis] = t;
                const size_t o = ((size_t)c[2] * ny + c[1]) * nx + c[0];
                fprintf(f, " %3d %3d %3d", c[0] + 1, c[1] + 1, c[2] + 1);
                for (int q = 0; q < 9; q++) fprintf(f, " % 9.3E", F[q][o]);
                fprintf(f, "\n");
            }
        }
        fclose(f);
    }
    float *own[9] = {hxSource, hySource, hzSource, dxSource, dySource, dzSource, exSource, eySource, ezSource};
    for (auto p : own) free(p);
}

// fdtd.cu:1288-1494: <out>/{H,D,E}_field_{x,y,z}_%05d_z%05d.out, ny lines of
// nx values " % 9.3E"; 0-based iteration and z.  Frees the nine buffers.
extern "C" void writeResults3d(FdtdParams *params, float *hxSource, float *hySource, float *hzSource,
                               float *dxSource, float *dySource, float *dzSource, float *exSource,
                               float *eySource, float *ezSource, int currentIteration)
{
    const int nx = params->nx, ny = params->ny, nz = params->nz;
    const char kind[3] = {'H', 'D', 'E'};
    float *F[9] = {hxSource, hySource, hzSource, dxSource, dySource, dzSource, exSource, eySource, ezSource};
    char path[1100];
    std::string line;
    for (int fidx = 0; fidx < 9; fidx++) {
        for (int iz = 0; iz < nz; iz++) {
            snprintf(path, sizeof path, "%s/%c_field_%c_%05d_z%05d.out", params->outputPath, kind[fidx / 3],
                     "xyz"[fidx % 3], currentIteration, iz);
            FILE *f = open_out(path);
            for (int iy = 0; iy < ny; iy++) {
                const float *row = F[fidx] + ((size_t)iz * ny + iy) * nx;
                for (int ix = 0; ix < nx; ix++) fprintf(f, " % 9.3E", row[ix]);
                fprintf(f, "\n");
            }
            fclose(f);
        }
    }
    for (auto p : F) free(p);
}
