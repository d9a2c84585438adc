// fdtd -- command-line driver with the reference's inputs and outputs
// (fdtd.cu:46-544): reads the parameter file and tissue table, loads the
// phantom slices, sets up Mur planes (from the default material, before the
// materials load, as fdtd.cu:59-65 does) and the source waveform, steps the
// grid and writes the text results.
//
//   fdtd [params] [--specs FILE] [--legacy] [--debye fixed|ref]
//        [--jz hold3|step] [--jz-limit N] [--every K] [--lines] [--quiet]
//
// Defaults follow the reference: data/input_params, data/mat_specs_riken,
// jz[3*floor(n/3)] limited to its first 512 entries, and all z slices of all
// nine fields written after every step (writeResults3d).  --lines writes the
// per-source line files of writeResults instead.  --legacy runs the five
// drop-in entry points on the reference's FdtdField layout; the default runs
// the fused context path.  The Debye update defaults to the stable form
// (d = a + b - sigma dt^2); --debye ref selects fdtd_calculations.cu:144.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fdtd.h"
#include "fdtd_amd.h"
#include "fdtd_calculations.h"

#define CHECK(call)                                                           \
    do {                                                                      \
        const hipError_t error_ = (call);                                     \
        if (error_ != hipSuccess) {                                           \
            printf("Error: %s: %d\n", __FILE__, __LINE__);                    \
            printf("Code %d: %s\n", (int)error_, hipGetErrorString(error_));  \
            exit(EXIT_FAILURE);                                               \
        }                                                                     \
    } while (0)

#define CHECK_RC(call)                                                        \
    do {                                                                      \
        const int rc_ = (call);                                               \
        if (rc_ != 0) {                                                       \
            printf("Error: %s: %d\n", __FILE__, __LINE__);                    \
            printf("Code %d: %s\n", rc_, fdtd_last_error());                  \
            exit(EXIT_FAILURE);                                               \
        }                                                                     \
    } while (0)

namespace {

struct Options {
    std::string params = "data/input_params";
    std::string specs = "data/mat_specs_riken";
    bool legacy = false;
    int debye = FDTD_DEBYE_FIXED;
    int jz_mode = FDTD_JZ_HOLD3;
    int jz_limit = D_MAX_JZ;
    int every = 1;
    bool lines = false;
    bool quiet = false;
};

[[noreturn]] void usage()
{
    printf("usage: fdtd [params] [--specs FILE] [--legacy] [--debye fixed|ref] [--jz hold3|step] "
           "[--jz-limit N] [--every K] [--lines] [--quiet]\n");
    exit(EXIT_FAILURE);
}

Options parse(int argc, char **argv)
{
    Options o;
    for (int a = 1; a < argc; a++) {
        std::string s = argv[a];
        auto next = [&]() -> std::string { if (a + 1 >= argc) usage(); return argv[++a]; };
        if (s == "--specs") o.specs = next();
        else if (s == "--legacy") o.legacy = true;
        else if (s == "--debye") { std::string v = next(); o.debye = v == "ref" ? FDTD_DEBYE_REF : FDTD_DEBYE_FIXED; }
        else if (s == "--jz") { std::string v = next(); o.jz_mode = v == "step" ? FDTD_JZ_STEP : FDTD_JZ_HOLD3; }
        else if (s == "--jz-limit") o.jz_limit = atoi(next().c_str());
        else if (s == "--every") o.every = atoi(next().c_str());
        else if (s == "--lines") o.lines = true;
        else if (s == "--quiet") o.quiet = true;
        else if (s.size() > 1 && s[0] == '-') usage();
        else o.params = s;
    }
    if (o.every < 1) o.every = 1;
    return o;
}

// Writes one set of results from nine host buffers (malloc'ed; the writers
// free them, as fdtd.cu:1274-1284 and 1481-1493 do).
void write_set(FdtdParams *p, float **f, int iteration, bool lines)
{
    if (lines)
        writeResults(p, f[3], f[4], f[5], f[0], f[1], f[2], f[6], f[7], f[8], iteration);
    else
        writeResults3d(p, f[3], f[4], f[5], f[0], f[1], f[2], f[6], f[7], f[8], iteration);
}

}  // namespace

int main(int argc, char **argv)
{
    const Options opt = parse(argc, argv);
    printf("Reading parameters...\n");
    FdtdParams *params = initParamsWithPath(opt.params.c_str());
    if (!opt.quiet) printParams(params);

    printf("Initializing field...\n");
    FdtdField *field = initFieldWithParams(params);
    setupMurBoundary(params, field);

    printf("Reading materials data...\n");
    loadMaterials(params, field, opt.specs.c_str(), params->inputPath);

    printf("Initializing sources...\n");
    setupSources(params);

    const size_t n = (size_t)params->nx * params->ny * params->nz;
    const size_t bytes = n * sizeof(float);
    hipStream_t stream;
    CHECK(hipStreamCreate(&stream));
    std::vector<float *> host(9);
    auto take = [&](const std::vector<const float *> &dev) {
        for (int f = 0; f < 9; f++) {
            host[f] = (float *)malloc(bytes);
            if (!host[f]) { printf("Couldn't allocate %zu bytes\n", bytes); exit(EXIT_FAILURE); }
            CHECK(hipMemcpyAsync(host[f], dev[f], bytes, hipMemcpyDeviceToHost, stream));
        }
        CHECK(hipStreamSynchronize(stream));
    };

    if (opt.legacy) {
        printf("Initializing device field...\n");
        FdtdField *dev = initDeviceFieldWithParams(params);
        printf("Copying data to GPU...\n\n");
        copyDataToDevice(params, field, dev);
        int jzc = params->jzCount;
        params->jzCount = opt.jz_limit < jzc ? opt.jz_limit : jzc;
        CHECK_RC(copySymbolsToDevice(params));
        params->jzCount = jzc;
        CHECK_RC(fdtdSetDebyeMode(opt.debye));
        float *E[3][3] = {{dev->ex0, dev->ey0, dev->ez0}, {dev->ex1, dev->ey1, dev->ez1}, {dev->ex2, dev->ey2, dev->ez2}};
        float *D[3][3] = {{dev->dx0, dev->dy0, dev->dz0}, {dev->dx1, dev->dy1, dev->dz1}, {dev->dx2, dev->dy2, dev->dz2}};
        for (int it = 0; it < params->iterationsCount; it++) {
            if (!opt.quiet) printf("Running iteration %d\n", it + 1);
            const int t = it % 3, s1 = (it + 2) % 3, s2 = (it + 1) % 3;
            const int jzi = opt.jz_mode == FDTD_JZ_HOLD3 ? (it / 3) * 3 : it;
            CHECK_RC(updateHField(dev->hx, dev->hy, dev->hz, E[s1][0], E[s1][1], E[s1][2], stream));
            CHECK_RC(updateDField(D[t][0], D[t][1], D[t][2], D[s1][0], D[s1][1], D[s1][2], dev->hx, dev->hy, dev->hz, stream));
            CHECK_RC(updateSources(D[t][2], D[s1][2], dev->hx, dev->hy, jzi, stream));
            CHECK_RC(updateEField(E[t][0], E[t][1], E[t][2], E[s1][0], E[s1][1], E[s1][2], E[s2][0], E[s2][1], E[s2][2],
                                  D[t][0], D[t][1], D[t][2], D[s1][0], D[s1][1], D[s1][2], D[s2][0], D[s2][1], D[s2][2],
                                  dev->sigma, dev->epsI, dev->epsS, dev->tauD, stream));
            CHECK_RC(updateMurBoundary(E[t][0], E[t][1], E[t][2], E[s1][0], E[s1][1], E[s1][2], dev->rpx0, dev->rpy0,
                                       dev->rpz0, dev->rpxEnd, dev->rpyEnd, dev->rpzEnd, stream));
            if ((it + 1) % opt.every == 0) {
                take({D[t][0], D[t][1], D[t][2], dev->hx, dev->hy, dev->hz, E[t][0], E[t][1], E[t][2]});
                write_set(params, host.data(), it, opt.lines);
            }
        }
        CHECK(hipStreamSynchronize(stream));
        deallocDeviceField(dev);
    } else {
        FdtdExt ext;
        fdtd_ext_defaults(&ext);
        ext.debye_mode = opt.debye;
        ext.jz_index_mode = opt.jz_mode;
        ext.jz_limit = opt.jz_limit;
        fdtd_ctx *ctx = nullptr;
        CHECK_RC(fdtd_ctx_create(params, &ext, &ctx));
        CHECK_RC(fdtd_ctx_set_material_arrays(ctx, field->sigma, field->epsS, field->epsI, field->tauD));
        CHECK_RC(fdtd_ctx_set_rp(ctx, field->rpx0, field->rpy0, field->rpz0, field->rpxEnd, field->rpyEnd, field->rpzEnd));
        CHECK_RC(fdtd_ctx_set_sources(ctx, params->sourcesCount, params->sources, params->jzCount, params->jz));
        for (int it = 0; it < params->iterationsCount; it++) {
            if (!opt.quiet) printf("Running iteration %d\n", it + 1);
            CHECK_RC(fdtd_step_fused(ctx, stream));
            if ((it + 1) % opt.every == 0) {
                for (int f = 0; f < 9; f++) {
                    host[f] = (float *)malloc(bytes);
                    if (!host[f]) { printf("Couldn't allocate %zu bytes\n", bytes); exit(EXIT_FAILURE); }
                    CHECK_RC(fdtd_ctx_get_field(ctx, f, 0, host[f]));
                }
                write_set(params, host.data(), it, opt.lines);
            }
        }
        CHECK(hipStreamSynchronize(stream));
        fdtd_ctx_destroy(ctx);
    }
    CHECK(hipStreamDestroy(stream));
    deallocField(field);
    deallocParams(params);
    return 0;
}
