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
//
// Result output overlaps the steps, as the reference's copy and writer
// threads do (fdtd.cu:17-43, 134-155, 253, 1136-1194); see Pipeline below.
// --writers N bounds the dumps being written at once (default 7, the
// reference's MAX_COPY_THREADS); --sync writes each dump on the stepping
// thread before the next step instead.
#include <hip/hip_runtime.h>
#include <pthread.h>

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
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
    bool sync = false;
    int writers = 7;
};

[[noreturn]] void usage()
{
    printf("usage: fdtd [params] [--specs FILE] [--legacy] [--debye fixed|ref] [--jz hold3|step] "
           "[--jz-limit N] [--every K] [--lines] [--quiet] [--writers N] [--sync]\n");
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
        else if (s == "--sync") o.sync = true;
        else if (s == "--writers") o.writers = atoi(next().c_str());
        else if (s.size() > 1 && s[0] == '-') usage();
        else o.params = s;
    }
    if (o.every < 1) o.every = 1;
    if (o.writers < 1) o.writers = 1;
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

// writeResultsWithParams (fdtd.cu:1163-1194) with the line writer: waits for
// the three copy threads, writes, frees what they allocated.
void *write_lines_with_params(void *arg)
{
    ResultsParams *rp = (ResultsParams *)arg;
    CopyParams *c[3] = {rp->hParams, rp->dParams, rp->eParams};
    for (CopyParams *q : c) pthread_mutex_lock(q->copyMutex);
    writeResults(rp->params, c[0]->xBuffer, c[0]->yBuffer, c[0]->zBuffer, c[1]->xBuffer, c[1]->yBuffer,
                 c[1]->zBuffer, c[2]->xBuffer, c[2]->yBuffer, c[2]->zBuffer, rp->currentIteration);
    for (CopyParams *q : c) {
        pthread_mutex_unlock(q->copyMutex);
        free(q->copyMutex);
        free(q);
    }
    free(rp);
    return nullptr;
}

// Result output overlapped with the steps.  Per dump:
//   1. the nine newest fields are copied on the compute stream into a device
//      staging set (fdtd_ctx_snapshot, or nine device copies on the legacy
//      path), so later steps may overwrite the ring slots at once;
//   2. the set's own side stream waits for that copy and moves the set to
//      pinned host memory in one transfer;
//   3. three copy threads (H, D, E) run copyResultsWithParams: each waits for
//      the side stream and copies its three volumes out of pinned memory
//      into malloc'ed buffers, after which the set can be reused;
//   4. a writer thread runs writeResultsWithParams (or the line writer) on
//      those buffers, which frees them.
// At most `writers` dumps are in flight (the reference's MAX_COPY_THREADS
// throttle, fdtd.cu:17-43, here a condition variable instead of a spin);
// the stepping thread waits only for that throttle or for a staging set
// still being copied out.  The copy mutexes are locked here and unlocked by
// the copy threads, as the reference's pipeline does (fdtd.cu:151-155,
// 1159): glibc's default mutex allows that.
class Pipeline {
public:
    Pipeline(FdtdParams *p, bool lines, int nsets, int writers)
        : p_(p), lines_(lines), writers_max_(writers), n_((size_t)p->nx * p->ny * p->nz)
    {
        sets_.resize(nsets);
        for (Set &s : sets_) {
            CHECK(hipMalloc(&s.dev, 9 * n_ * sizeof(float)));
            CHECK(hipHostMalloc(&s.pinned, 9 * n_ * sizeof(float), hipHostMallocDefault));
            CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
            CHECK(hipEventCreateWithFlags(&s.ready, hipEventDisableTiming));
        }
    }

    ~Pipeline()
    {
        drain();
        for (Set &s : sets_) {
            (void)hipFree(s.dev);
            (void)hipHostFree(s.pinned);
            (void)hipStreamDestroy(s.stream);
            (void)hipEventDestroy(s.ready);
        }
    }

    // The device staging buffer of the next dump (9 volumes, field order Dx
    // Dy Dz Hx Hy Hz Ex Ey Ez), once its previous contents are out and a
    // writer slot is free.
    float *begin()
    {
        const auto t0 = std::chrono::steady_clock::now();
        Set &s = sets_[next_];
        join_copies(s);
        {
            std::unique_lock<std::mutex> lk(m_);
            cv_.wait(lk, [&] { return writers_ < writers_max_; });
            writers_++;
        }
        wait_s_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return s.dev;
    }

    // Queues the transfer of the set begin() returned, after the work queued
    // on `compute` so far, and starts its copy and writer threads.
    void commit(hipStream_t compute, int iteration)
    {
        Set &s = sets_[next_];
        next_ = (next_ + 1) % (int)sets_.size();
        CHECK(hipEventRecord(s.ready, compute));
        CHECK(hipStreamWaitEvent(s.stream, s.ready, 0));
        CHECK(hipMemcpyAsync(s.pinned, s.dev, 9 * n_ * sizeof(float), hipMemcpyDeviceToHost, s.stream));
        CopyParams *cp[3];
        for (int g = 0; g < 3; g++) {  // H, D, E: fields 3-5, 0-2, 6-8 of the set
            const int f0 = g == 0 ? 3 : (g == 1 ? 0 : 6);
            cp[g] = (CopyParams *)calloc(1, sizeof(CopyParams));
            if (!cp[g]) { printf("Couldn't allocate copy parameters\n"); exit(EXIT_FAILURE); }
            cp[g]->xSource = s.pinned + (size_t)f0 * n_;
            cp[g]->ySource = s.pinned + (size_t)(f0 + 1) * n_;
            cp[g]->zSource = s.pinned + (size_t)(f0 + 2) * n_;
            cp[g]->params = p_;
            cp[g]->stream = s.stream;
            cp[g]->copyMutex = (pthread_mutex_t *)malloc(sizeof(pthread_mutex_t));
            if (!cp[g]->copyMutex) { printf("Couldn't allocate a mutex\n"); exit(EXIT_FAILURE); }
            pthread_mutex_init(cp[g]->copyMutex, nullptr);
            pthread_mutex_lock(cp[g]->copyMutex);
            if (pthread_create(&s.copier[g], nullptr, copyResultsWithParams, cp[g]) != 0) {
                printf("Couldn't start a copy thread\n");
                exit(EXIT_FAILURE);
            }
        }
        s.copying = true;
        ResultsParams *rp = (ResultsParams *)calloc(1, sizeof(ResultsParams));
        Job *job = new Job{this, rp};
        if (!rp) { printf("Couldn't allocate result parameters\n"); exit(EXIT_FAILURE); }
        rp->params = p_;
        rp->hParams = cp[0];
        rp->dParams = cp[1];
        rp->eParams = cp[2];
        rp->currentIteration = iteration;
        pthread_t w;
        if (pthread_create(&w, nullptr, &Pipeline::writer_main, job) != 0) {
            printf("Couldn't start a writer thread\n");
            exit(EXIT_FAILURE);
        }
        pthread_detach(w);
    }

    // Waits until every dump is written.
    void drain()
    {
        for (Set &s : sets_) join_copies(s);
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return writers_ == 0; });
    }

    double wait_seconds() const { return wait_s_; }

private:
    struct Set {
        float *dev = nullptr, *pinned = nullptr;
        hipStream_t stream = nullptr;
        hipEvent_t ready = nullptr;
        pthread_t copier[3];
        bool copying = false;
    };
    struct Job {
        Pipeline *self;
        ResultsParams *rp;
    };

    static void join_copies(Set &s)
    {
        if (!s.copying) return;
        for (pthread_t &t : s.copier) pthread_join(t, nullptr);
        s.copying = false;
    }

    static void *writer_main(void *arg)
    {
        Job *job = (Job *)arg;
        Pipeline *self = job->self;
        if (self->lines_) write_lines_with_params(job->rp);
        else writeResultsWithParams(job->rp);
        delete job;
        {
            std::lock_guard<std::mutex> lk(self->m_);
            self->writers_--;
        }
        self->cv_.notify_all();
        return nullptr;
    }

    FdtdParams *p_;
    bool lines_;
    int writers_max_;
    size_t n_;
    std::vector<Set> sets_;
    int next_ = 0;
    std::mutex m_;
    std::condition_variable cv_;
    int writers_ = 0;
    double wait_s_ = 0.0;
};

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
    // --sync: the nine volumes to host on the stepping thread, then the text, then the next step
    std::vector<float *> host(9);
    auto take = [&](const std::vector<const float *> &dev) {
        for (int f = 0; f < 9; f++) {
            host[f] = (float *)malloc(bytes);
            if (!host[f]) { printf("Couldn't allocate %zu bytes\n", bytes); exit(EXIT_FAILURE); }
            CHECK(hipMemcpyAsync(host[f], dev[f], bytes, hipMemcpyDeviceToHost, stream));
        }
        CHECK(hipStreamSynchronize(stream));
    };
    // two staging sets: one filled on the device while the other is copied out
    const bool dumps = params->iterationsCount >= opt.every;
    Pipeline *pipe = (!opt.sync && dumps) ? new Pipeline(params, opt.lines, 2, opt.writers) : nullptr;
    const auto t_start = std::chrono::steady_clock::now();

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
                const float *src[9] = {D[t][0], D[t][1], D[t][2], dev->hx, dev->hy, dev->hz, E[t][0], E[t][1], E[t][2]};
                if (pipe) {
                    float *stage = pipe->begin();
                    for (int f = 0; f < 9; f++)
                        CHECK(hipMemcpyAsync(stage + f * n, src[f], bytes, hipMemcpyDeviceToDevice, stream));
                    pipe->commit(stream, it);
                } else {
                    take({src[0], src[1], src[2], src[3], src[4], src[5], src[6], src[7], src[8]});
                    write_set(params, host.data(), it, opt.lines);
                }
            }
        }
        CHECK(hipStreamSynchronize(stream));
        if (pipe) pipe->drain();
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
                if (pipe) {
                    float *stage = pipe->begin();
                    CHECK_RC(fdtd_ctx_snapshot(ctx, stage, stream));
                    pipe->commit(stream, it);
                } else {
                    for (int f = 0; f < 9; f++) {
                        host[f] = (float *)malloc(bytes);
                        if (!host[f]) { printf("Couldn't allocate %zu bytes\n", bytes); exit(EXIT_FAILURE); }
                        CHECK_RC(fdtd_ctx_get_field(ctx, f, 0, host[f]));
                    }
                    write_set(params, host.data(), it, opt.lines);
                }
            }
        }
        CHECK(hipStreamSynchronize(stream));
        if (pipe) pipe->drain();
        fdtd_ctx_destroy(ctx);
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (!opt.quiet || getenv("FDTD_TIMING"))
        printf("Stepping and output: %.3f s; the stepping thread waited %.3f s for the result writers\n", wall,
               pipe ? pipe->wait_seconds() : 0.0);
    delete pipe;
    CHECK(hipStreamDestroy(stream));
    deallocField(field);
    deallocParams(params);
    return 0;
}
