// fdtd_dropin_solve without device allocation: the grid's state lives in a
// file-scope device array (`scratch`, kDropinScratchBytes per GPU, held until
// the process ends) instead of hipMalloc'ed buffers, for callers that cannot
// allocate (code under HIP graph capture, serving loops that forbid it).
//
// The host never passes `scratch` itself to a kernel or a copy (in host code
// the name is a stand-in, not a device address): on the first call on a
// device, with that device current, hipGetSymbolAddress gives the device's
// own address of its copy, which is kept per device and used everywhere.
// Every call on one GPU shares that GPU's array and leaves the last call's
// contents in it, so the state is cleared (fields to zero) at the start of
// each call by a fill kernel, and calls that overlap on one GPU race (as
// fdtd_dropin_solve's calls already must not overlap: both use the
// process-global symbols of copySymbolsToDevice).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "fdtd_amd.h"
#include "fdtd_internal.h"

namespace {

constexpr size_t kDropinScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kDropinScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices];   // device address of `scratch`, per device

__global__ __launch_bounds__(256) void zero_words(uint4 *p, long long words)
{
    for (long long w = (long long)blockIdx.x * 256 + threadIdx.x; w < words; w += (long long)gridDim.x * 256)
        p[w] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ __launch_bounds__(256) void fill_ones(float *p, long long n)
{
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) p[t] = 1.0f;
}

// The four per-cell material arrays from uint8 ids and the table rows
// {sigma, eps_s, eps_inf, tau} (ids null: material 0 everywhere), as
// fdtd_dropin_solve expands them on the host.
__global__ __launch_bounds__(256) void expand_materials(const unsigned char *ids, const float *table, long long n,
                                                        float *sigma, float *eps_s, float *eps_i, float *tau)
{
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
        const int m = ids ? ids[t] : 0;
        sigma[t] = table[4 * m + 0];
        eps_s[t] = table[4 * m + 1];
        eps_i[t] = table[4 * m + 2];
        tau[t] = table[4 * m + 3];
    }
}

unsigned grid_for(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 2048); }

// Carves 256-byte aligned pieces off the scratch array.
struct Carve {
    char *base;
    size_t used = 0;
    template <typename T> T *take(size_t count)
    {
        T *p = (T *)(base + used);
        used += (count * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" int fdtd_dropin_solve_noalloc(const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *in,
                                         double *probes_host, hipStream_t stream)
{
    FdtdExt e;
    if (int rc = dropin_check(p, ext, in, probes_host, &e)) return rc;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const size_t N = (size_t)nx * ny * nz;
    const size_t rps[6] = {(size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2,
                           (size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2};
    // bytes the state takes: 21 field and 4 material arrays, the rp planes, the ids and the table, the probes
    size_t need = 25 * round256(N * sizeof(float)) + round256(N) + round256(sizeof(float) * 4 * in->nmat);
    for (size_t r : rps) need += round256(r * sizeof(float));
    need += round256(sizeof(int) * 4 * (size_t)in->nprobe) + round256(sizeof(double) * (size_t)in->nprobe * in->steps);
    if (need > kDropinScratchBytes)
        return fdtd_fail(FDTD_ERANGE,
                         "grid %dx%dx%d with %d probes over %d steps needs %zu bytes of state; the allocation-free "
                         "solve holds at most %zu (about 640 K cells)",
                         nx, ny, nz, in->nprobe, in->steps, need, kDropinScratchBytes);

    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fdtd_fail(FDTD_ERANGE, "device %d", dev);
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_scratch[dev]) {
            void *ptr = nullptr;
            const hipError_t he = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (he != hipSuccess) return fdtd_hip_fail(he, "hipGetSymbolAddress(scratch)");
            g_scratch[dev] = ptr;
        }
        base = g_scratch[dev];
    }

    Carve cv{(char *)base};
    DropinState st;
    for (int s = 0; s < 3; s++)
        for (int q = 0; q < 3; q++) st.E[s][q] = cv.take<float>(N);
    for (int s = 0; s < 3; s++)
        for (int q = 0; q < 3; q++) st.D[s][q] = cv.take<float>(N);
    for (int q = 0; q < 3; q++) st.H[q] = cv.take<float>(N);
    const size_t field_bytes = cv.used;   // the 21 field arrays start at zero, like fdtd_dropin_solve's
    for (int c = 0; c < 4; c++) st.mat[c] = cv.take<float>(N);
    for (int r = 0; r < 6; r++) st.rp[r] = cv.take<float>(rps[r]);
    unsigned char *ids = in->ids ? cv.take<unsigned char>(N) : nullptr;
    float *table = cv.take<float>(4 * (size_t)in->nmat);
    if (in->nprobe > 0) {
        st.probes = cv.take<int>(4 * (size_t)in->nprobe);
        st.probe_out = cv.take<double>((size_t)in->nprobe * in->steps);
    }

    zero_words<<<grid_for((long long)(field_bytes / 16)), 256, 0, stream>>>((uint4 *)base, (long long)(field_bytes / 16));
    HIP_TRY(hipGetLastError());
    if (ids) HIP_TRY(hipMemcpyAsync(ids, in->ids, N, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(table, in->table, sizeof(float) * 4 * in->nmat, hipMemcpyHostToDevice, stream));
    expand_materials<<<grid_for((long long)N), 256, 0, stream>>>(ids, table, (long long)N, st.mat[0], st.mat[1],
                                                                 st.mat[2], st.mat[3]);
    HIP_TRY(hipGetLastError());
    for (int r = 0; r < 6; r++) {
        if (in->rp[r]) {
            HIP_TRY(hipMemcpyAsync(st.rp[r], in->rp[r], rps[r] * sizeof(float), hipMemcpyHostToDevice, stream));
        } else {
            fill_ones<<<grid_for((long long)rps[r]), 256, 0, stream>>>(st.rp[r], (long long)rps[r]);
            HIP_TRY(hipGetLastError());
        }
    }
    if (in->nprobe > 0)
        HIP_TRY(hipMemcpyAsync(st.probes, in->probes, sizeof(int) * 4 * in->nprobe, hipMemcpyHostToDevice, stream));
    // the pageable host inputs are read by the copies above before this returns
    HIP_TRY(hipStreamSynchronize(stream));
    return dropin_run(p, e, in, st, probes_host, stream);
}
