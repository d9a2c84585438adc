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

int dropin_check(const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *in, const double *probes_host,
                 FdtdExt *out)
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
    if (in->ids) {
        const size_t N = (size_t)nx * ny * nz;
        for (size_t t = 0; t < N; t++)
            if (in->ids[t] >= in->nmat) return fdtd_fail(FDTD_ERANGE, "material id %d >= nmat", in->ids[t]);
    }
    *out = e;
    return 0;
}

int dropin_run(const FdtdParams *p, const FdtdExt &e, const FdtdInstance *in, const DropinState &st,
               double *probes_host, hipStream_t stream)
{
    const int nx = p->nx, ny = p->ny;
    // copySymbolsToDevice reads the scalars, sources and jz of the params (it only reads them).
    FdtdParams q = *p;
    q.sourcesCount = in->nsrc;
    q.sources = const_cast<int *>(in->src_xyz);
    const int lim = e.jz_limit > 0 ? std::min(e.jz_limit, in->jz_len) : in->jz_len;
    q.jzCount = lim;
    q.jz = const_cast<float *>(in->jz);
    if (int rc = copySymbolsToDevice(&q)) return rc;
    if (int rc = fdtdSetDebyeMode(e.debye_mode)) return rc;

    float *const(*E)[3] = st.E, *const(*D)[3] = st.D;
    const float *const *H = st.H, *const *mat = st.mat, *const *rp = st.rp;
    ProbeArgs pa;
    for (int f = 0; f < 9; f++) pa.factor[f] = 1.0;
    pa.probes = st.probes;
    pa.nprobe = in->nprobe;
    pa.pitch = nx;
    pa.plane = (long long)nx * ny;
    pa.out = st.probe_out;
    float *h0 = const_cast<float *>(H[0]), *h1 = const_cast<float *>(H[1]), *h2 = const_cast<float *>(H[2]);
    auto m = [&](int c) { return const_cast<float *>(mat[c]); };
    auto r = [&](int c) { return const_cast<float *>(rp[c]); };
    for (int it = 0; it < in->steps; it++) {
        const int t = it % 3, s1 = (it + 2) % 3, s2 = (it + 1) % 3;
        const int jzi = e.jz_index_mode == FDTD_JZ_HOLD3 ? (it / 3) * 3 : it;
        int rc = updateHField(h0, h1, h2, E[s1][0], E[s1][1], E[s1][2], stream);
        if (!rc) rc = updateDField(D[t][0], D[t][1], D[t][2], D[s1][0], D[s1][1], D[s1][2], h0, h1, h2, stream);
        if (!rc) rc = updateSources(D[t][2], D[s1][2], h0, h1, jzi, stream);
        if (!rc)
            rc = updateEField(E[t][0], E[t][1], E[t][2], E[s1][0], E[s1][1], E[s1][2], E[s2][0], E[s2][1], E[s2][2],
                              D[t][0], D[t][1], D[t][2], D[s1][0], D[s1][1], D[s1][2], D[s2][0], D[s2][1], D[s2][2],
                              m(0), m(2), m(1), m(3), stream);
        if (!rc) rc = updateMurBoundary(E[t][0], E[t][1], E[t][2], E[s1][0], E[s1][1], E[s1][2], r(0), r(1), r(2), r(3),
                                        r(4), r(5), stream);
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
        HIP_TRY(hipMemcpyAsync(probes_host, st.probe_out, sizeof(double) * (size_t)in->nprobe * in->steps,
                               hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
}

extern "C" int fdtd_dropin_solve(const FdtdParams *p, const FdtdExt *ext, const FdtdInstance *in, double *probes_host,
                                 hipStream_t stream)
{
    FdtdExt e;
    if (int rc = dropin_check(p, ext, in, probes_host, &e)) return rc;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const size_t N = (size_t)nx * ny * nz;

    // FdtdField-style state (fdtd.cu:764-872): E and D rings, H, four material
    // arrays, six rp planes.
    DevBufs B;
    DropinState st;
    const size_t rps[6] = {(size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2,
                           (size_t)2 * ny * nz, (size_t)nx * 2 * nz, (size_t)nx * ny * 2};
    for (int s = 0; s < 3; s++)
        for (int q = 0; q < 3; q++) {
            HIP_TRY(B.get(&st.E[s][q], N));
            HIP_TRY(B.get(&st.D[s][q], N));
        }
    for (int q = 0; q < 3; q++) HIP_TRY(B.get(&st.H[q], N));
    std::vector<float> host(N);
    for (int c = 0; c < 4; c++) {
        HIP_TRY(B.get(&st.mat[c], N));
        for (size_t t = 0; t < N; t++) {
            const int m = in->ids ? in->ids[t] : 0;
            host[t] = in->table[4 * m + c];   // sigma, eps_s, eps_inf, tau
        }
        HIP_TRY(hipMemcpy(st.mat[c], host.data(), N * sizeof(float), hipMemcpyHostToDevice));
    }
    for (int r = 0; r < 6; r++) {
        HIP_TRY(B.get(&st.rp[r], rps[r]));
        std::vector<float> ones;
        const float *src = in->rp[r];
        if (!src) {
            ones.assign(rps[r], 1.0f);
            src = ones.data();
        }
        HIP_TRY(hipMemcpy(st.rp[r], src, rps[r] * sizeof(float), hipMemcpyHostToDevice));
    }
    if (in->nprobe > 0) {
        HIP_TRY(hipMalloc(&st.probes, sizeof(int) * 4 * in->nprobe));
        B.all.push_back(st.probes);
        HIP_TRY(hipMemcpy(st.probes, in->probes, sizeof(int) * 4 * in->nprobe, hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&st.probe_out, sizeof(double) * (size_t)in->nprobe * in->steps));
        B.all.push_back(st.probe_out);
    }
    return dropin_run(p, e, in, st, probes_host, stream);
}
