// One grid solved through the five drop-in entry points, the way the
// reference's main loop drives them (fdtd.cu:59-72 set-up, the step loop of
// fdtd.cu:115-530: H -> D -> sources -> E -> Mur with the ring-slot
// rotation), with the per-voxel material arrays of FdtdField and the
// reference's limits (16 sources, 512 jz values).  Probe rows are recorded on
// the device after every step.  It allocates and frees the grid's device
// state itself and uses the process-global symbols of copySymbolsToDevice,
// so calls must not overlap.  This is the sequential baseline that batched
// and served runs are measured against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "fdtd_calculations.h"
#include "fdtd_internal.h"

namespace {

struct DevBufs {
    std::vector<void *> all;
    ~DevBufs()
    {
        for (void *p : all) (void)hipFree(p);
    }
    hipError_t get(float **p, size_t n)
    {
        hipError_t e = hipMalloc((void **)p, n * sizeof(float));
        if (e != hipSuccess) return e;
        all.push_back(*p);
        return hipMemset(*p, 0, n * sizeof(float));
    }
};

}  // namespace

extern "C" int fdtd_dropin_solve(const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *in, double *probes_host,
                                 hipStream_t stream)
{
    if (!p || !in) return fdtd_fail(FDTD_EINVAL, "null argument");
    FdtdExt e;
    fdtd_ext_defaults(&e);
    if (ext) e = *ext;
    if (int rc = fdtd_check_ext(e)) return rc;
    if (e.storage != FDTD_STORE_F32) return fdtd_fail(FDTD_EINVAL, "the drop-in path stores fp32 only");
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    if (nx < 3 || ny < 3 || nz < 3) return fdtd_fail(FDTD_ERANGE, "grid %dx%dx%d", nx, ny, nz);
    if (in->nmat < 1 || in->nmat > kMaxMaterials || !in->table || in->steps < 1) return fdtd_fail(FDTD_EINVAL, "bad grid inputs");
    if (in->nsrc < 0 || in->nsrc > D_MAX_SOURCES || (in->nsrc > 0 && !in->src_xyz))
        return fdtd_fail(FDTD_ERANGE, "the drop-in path takes at most %d sources", D_MAX_SOURCES);
    if (in->jz_len < 0 || (in->jz_len > 0 && !in->jz)) return fdtd_fail(FDTD_EINVAL, "bad waveform");
    if (in->nprobe < 0 || (in->nprobe > 0 && (!in->probes || !probes_host))) return fdtd_fail(FDTD_EINVAL, "bad probes");
    if (int rc = fdtd_check_sources(nx, ny, nz, in->nsrc, in->src_xyz)) return rc;
    if (int rc = fdtd_check_probes(nx, ny, nz, in->nprobe, in->probes)) return rc;
    const size_t N = (size_t)nx * ny * nz;

    // FdtdField-style state (fdtd.cu:764-872): E and D rings, H, four material
    // arrays, six rp planes.
    DevBufs B;
    float *E[3][3], *D[3][3], *H[3], *mat[4], *rp[6];
    const size_t rps[6] = {(size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2,
                           (size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2};
    for (int s = 0; s < 3; s++)
        for (int q = 0; q < 3; q++) {
            HIP_TRY(B.get(&E[s][q], N));
            HIP_TRY(B.get(&D[s][q], N));
        }
    for (int q = 0; q < 3; q++) HIP_TRY(B.get(&H[q], N));
    std::vector<float> host(N);
    for (int c = 0; c < 4; c++) {
        HIP_TRY(B.get(&mat[c], N));
        for (size_t t = 0; t < N; t++) {
            const int m = in->ids ? in->ids[t] : 0;
            if (m >= in->nmat) return fdtd_fail(FDTD_ERANGE, "material id %d >= nmat", m);
            host[t] = in->table[4 * m + c];   // sigma, eps_s, eps_inf, tau
        }
        HIP_TRY(hipMemcpy(mat[c], host.data(), N * sizeof(float), hipMemcpyHostToDevice));
    }
    for (int r = 0; r < 6; r++) {
        HIP_TRY(B.get(&rp[r], rps[r]));
        std::vector<float> ones;
        const float *src = in->rp[r];
        if (!src) {
            ones.assign(rps[r], 1.0f);
            src = ones.data();
        }
        HIP_TRY(hipMemcpy(rp[r], src, rps[r] * sizeof(float), hipMemcpyHostToDevice));
    }
    int *dprobes = nullptr;
    double *dout = nullptr;
    if (in->nprobe > 0) {
        HIP_TRY(hipMalloc(&dprobes, sizeof(int) * 4 * in->nprobe));
        B.all.push_back(dprobes);
        HIP_TRY(hipMemcpy(dprobes, in->probes, sizeof(int) * 4 * in->nprobe, hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&dout, sizeof(double) * (size_t)in->nprobe * in->steps));
        B.all.push_back(dout);
    }

    // copySymbolsToDevice reads the scalars, sources and jz of the params.
    FdtdParams q = *p;
    std::vector<int> src(in->src_xyz, in->src_xyz + 3 * in->nsrc);
    std::vector<float> jz(in->jz, in->jz + in->jz_len);
    q.sourcesCount = in->nsrc;
    q.sources = src.data();
    const int lim = e.jz_limit > 0 ? std::min(e.jz_limit, in->jz_len) : in->jz_len;
    q.jzCount = lim;
    q.jz = jz.data();
    if (int rc = copySymbolsToDevice(&q)) return rc;
    if (int rc = fdtdSetDebyeMode(e.debye_mode)) return rc;

    ProbeArgs pa;
    for (int f = 0; f < 9; f++) pa.factor[f] = 1.0;
    pa.probes = dprobes;
    pa.nprobe = in->nprobe;
    pa.pitch = nx;
    pa.plane = (long long)nx * ny;
    pa.out = dout;
    for (int it = 0; it < in->steps; it++) {
        const int t = it % 3, s1 = (it + 2) % 3, s2 = (it + 1) % 3;
        const int jzi = e.jz_index_mode == FDTD_JZ_HOLD3 ? (it / 3) * 3 : it;
        int rc = updateHField(H[0], H[1], H[2], E[s1][0], E[s1][1], E[s1][2], stream);
        if (!rc) rc = updateDField(D[t][0], D[t][1], D[t][2], D[s1][0], D[s1][1], D[s1][2], H[0], H[1], H[2], stream);
        if (!rc) rc = updateSources(D[t][2], D[s1][2], H[0], H[1], jzi, stream);
        if (!rc)
            rc = updateEField(E[t][0], E[t][1], E[t][2], E[s1][0], E[s1][1], E[s1][2], E[s2][0], E[s2][1], E[s2][2],
                              D[t][0], D[t][1], D[t][2], D[s1][0], D[s1][1], D[s1][2], D[s2][0], D[s2][1], D[s2][2],
                              mat[0], mat[2], mat[1], mat[3], stream);
        if (!rc)
            rc = updateMurBoundary(E[t][0], E[t][1], E[t][2], E[s1][0], E[s1][1], E[s1][2], rp[0], rp[1], rp[2], rp[3],
                                   rp[4], rp[5], stream);
        if (rc) return fdtd_hip_fail((hipError_t)rc, "drop-in step");
        if (in->nprobe > 0) {
            const void *fp[9] = {D[t][0], D[t][1], D[t][2], H[0], H[1], H[2], E[t][0], E[t][1], E[t][2]};
            for (int f = 0; f < 9; f++) pa.field[f] = fp[f];
            pa.row = it;
            const hipError_t he = launch_probes(pa, kStoreF32, stream);
            if (he != hipSuccess) return fdtd_hip_fail(he, "probe_cells");
        }
    }
    if (in->nprobe > 0)
        HIP_TRY(hipMemcpyAsync(probes_host, dout, sizeof(double) * (size_t)in->nprobe * in->steps, hipMemcpyDeviceToHost,
                               stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
}
