// Fused time-step kernels for gfx950.
//
// fused_step: one launch computes H^{n+1/2} (fdtd_calculations.cu:32-77),
// D^{n+1} from curl H (:80-122), the soft Jz correction (:187-213) and the
// Debye E^{n+1} (:125-184) for every cell, with the reference's ranges and
// rounding.  Layout and schedule:
//   * a block owns a 64 x TY column tile of the (x, y) plane and marches
//     `kchunk` z planes; lanes of a wave64 run along x, so every global
//     access of a wave is one contiguous 256-byte row segment;
//   * E^n of planes k and k+1 is staged in LDS with a one-cell halo (double
//     buffered), H^{n+1/2} of plane k in LDS with its -x/-y halo, and the
//     own-column H of plane k-1 in registers (the z-march);
//   * H is ping-ponged (read H^{n-1/2}, write H^{n+1/2}) because the halo
//     H is recomputed by neighbouring tiles, which an in-place update would
//     race with;
//   * material coefficients come from a per-material table gathered through
//     a uint8 id per cell (top bit = source cell), staged in LDS.
// mur_faces: the 12 (component, face) first-order Mur updates
// (fdtd_calculations.cu:216-428) over face cells only.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fdtd_internal.h"

namespace {

constexpr int TX = 64;

__device__ inline int xcd_logical(int b, int nblk)
{
    // Workgroups are dispatched to the 8 XCDs round-robin by id.  Give each
    // XCD a contiguous run of logical tiles so tiles that share halo planes
    // share an L2.
    const int q = nblk / 8, rem = nblk % 8;
    const int x = b % 8, s = b / 8;
    return x * q + min(x, rem) + s;
}

__device__ inline float e_update(float c, float d, float a, double g1, double g2, double g3, float e0, float e1,
                                 float dn1, float dn, float dm1)
{
    const float fe = d * e0 - a * e1;
    double s = (double)fe + g1 * (double)dn1;
    s = s - g2 * (double)dn;
    s = s + g3 * (double)dm1;
    return (float)((double)c * s);
}

__device__ __forceinline__ float ldg(const float *p, uint32_t boff)
{
    return *(const float *)((const char *)p + boff);
}

__device__ __forceinline__ void stg(float *p, uint32_t boff, float v) { *(float *)((char *)p + boff) = v; }

// All per-lane addresses are a 64-bit array base (uniform, in SGPRs) plus a
// 32-bit byte offset shared by every array indexed by the same cell, so the
// loads use the scalar-base form with one offset register.  The host keeps
// every array below 4 GiB.
template <int TY>
__global__ __launch_bounds__(TX *TY) void fused_step(const StepArgs A)
{
    __shared__ float sE[2][3][TY + 2][TX + 2];
    __shared__ float sH[3][TY + 1][TX + 1];
    __shared__ float lA[kMaxMaterials], lC[kMaxMaterials], lD[kMaxMaterials];
    __shared__ double lG1[kMaxMaterials], lG2[kMaxMaterials], lG3[kMaxMaterials];

    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const int nblk = A.tiles_x * A.tiles_y * A.tiles_z;
    const int b = A.xcd_remap ? xcd_logical((int)blockIdx.x, nblk) : (int)blockIdx.x;
    const int bx = b % A.tiles_x;
    const int by = (b / A.tiles_x) % A.tiles_y;
    const int bz = b / (A.tiles_x * A.tiles_y);

    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int i0 = bx * TX, j0 = by * TY;
    const int k0 = bz * A.kchunk;
    const int k1 = min(k0 + A.kchunk, nz);
    const int i = i0 + tx, j = j0 + ty;
    const bool own = (i < nx) && (j < ny);
    const uint32_t syb = (uint32_t)A.pitch * 4u, szb = (uint32_t)A.plane * 4u;
    const uint32_t colb = (uint32_t)j * syb + (uint32_t)i * 4u;  // byte offset of (i, j) within a plane

    for (int m = tid; m < A.nmat; m += TX * TY) {
        const MatCoef mc = A.lut[m];
        lA[m] = mc.a; lC[m] = mc.c; lD[m] = mc.d;
        lG1[m] = mc.g1; lG2[m] = mc.g2; lG3[m] = mc.g3;
    }

    auto ld = [&](const float *p, int ii, int jj, int kk) -> float {
        return (ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz)
                   ? ldg(p, (uint32_t)kk * szb + (uint32_t)jj * syb + (uint32_t)ii * 4u)
                   : 0.0f;
    };

    // Stage E^n of plane k (own cell plus the halo cells the H stencils of
    // this tile and of its -x/-y halo need) into buffer `buf`.
    auto load_plane = [&](int buf, int k) {
        const bool kin = (k >= 0 && k < nz);
        const uint32_t ob = (uint32_t)k * szb + colb;
        const bool ok = own && kin;
        sE[buf][0][ty + 1][tx + 1] = ok ? ldg(A.ex0, ob) : 0.0f;
        sE[buf][1][ty + 1][tx + 1] = ok ? ldg(A.ey0, ob) : 0.0f;
        sE[buf][2][ty + 1][tx + 1] = ok ? ldg(A.ez0, ob) : 0.0f;
        if (ty == 0) {
            sE[buf][0][0][tx + 1] = ld(A.ex0, i, j0 - 1, k);
            sE[buf][1][0][tx + 1] = ld(A.ey0, i, j0 - 1, k);
            sE[buf][2][0][tx + 1] = ld(A.ez0, i, j0 - 1, k);
        }
        if (ty == TY - 1) {
            sE[buf][0][TY + 1][tx + 1] = ld(A.ex0, i, j0 + TY, k);
            sE[buf][2][TY + 1][tx + 1] = ld(A.ez0, i, j0 + TY, k);
        }
        if (tx == 0) {
            sE[buf][0][ty + 1][0] = ld(A.ex0, i0 - 1, j, k);
            sE[buf][1][ty + 1][0] = ld(A.ey0, i0 - 1, j, k);
            sE[buf][2][ty + 1][0] = ld(A.ez0, i0 - 1, j, k);
            if (ty == TY - 1) sE[buf][0][TY + 1][0] = ld(A.ex0, i0 - 1, j0 + TY, k);
        }
        if (tx == TX - 1) {
            sE[buf][1][ty + 1][TX + 1] = ld(A.ey0, i0 + TX, j, k);
            sE[buf][2][ty + 1][TX + 1] = ld(A.ez0, i0 + TX, j, k);
            if (ty == 0) sE[buf][1][0][TX + 1] = ld(A.ey0, i0 + TX, j0 - 1, k);
        }
    };

    const float chx = A.chx, chy = A.chy, chz = A.chz;
    const float cdx = A.cdx, cdy = A.cdy, cdz = A.cdz;

    float hxm = 0.0f, hym = 0.0f;  // own-column Hx, Hy of plane k-1 (time level n+1/2)
    int cur = 0;
    if (k0 > 0) {
        load_plane(0, k0 - 1);
        load_plane(1, k0);
        __syncthreads();
        const int k = k0 - 1;
        if (own) {
            const uint32_t ob = (uint32_t)k * szb + colb;
            float hx = ldg(A.hxo, ob), hy = ldg(A.hyo, ob);
            const float exk = sE[0][0][ty + 1][tx + 1], eyk = sE[0][1][ty + 1][tx + 1], ezk = sE[0][2][ty + 1][tx + 1];
            const float exk1 = sE[1][0][ty + 1][tx + 1], eyk1 = sE[1][1][ty + 1][tx + 1];
            if (i >= 1 && i < nx - 1 && j < ny - 1 && k < nz - 1)
                hx = hx - chy * (sE[0][2][ty + 2][tx + 1] - ezk) + chy * (eyk1 - eyk);
            if (i < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1)
                hy = hy - chz * (exk1 - exk) + chx * (sE[0][2][ty + 1][tx + 2] - ezk);
            hxm = hx;
            hym = hy;
        }
        cur = 1;
        __syncthreads();
    } else {
        load_plane(0, k0);
    }

    const bool row_halo = (ty == 0) && (j0 >= 1) && (i < nx);
    const bool col_halo = (tx == 0) && (i0 >= 1) && (j < ny);
    const uint32_t rowb = (uint32_t)(j0 - 1) * syb + (uint32_t)i * 4u;   // (i, j0-1)
    const uint32_t colhb = (uint32_t)j * syb + (uint32_t)(i0 - 1) * 4u;  // (i0-1, j)

    for (int k = k0; k < k1; ++k) {
        const int nxt = cur ^ 1;
        const uint32_t ob = (uint32_t)k * szb + colb;

        // Own-cell operands of this plane.
        float hxo = 0.f, hyo = 0.f, hzo = 0.f, dxn = 0.f, dyn = 0.f, dzn = 0.f, dxm = 0.f, dym = 0.f, dzm = 0.f;
        float exm = 0.f, eym = 0.f, ezm = 0.f;
        int id = 0;
        if (own) {
            hxo = ldg(A.hxo, ob); hyo = ldg(A.hyo, ob); hzo = ldg(A.hzo, ob);
            dxn = ldg(A.dx0, ob); dyn = ldg(A.dy0, ob); dzn = ldg(A.dz0, ob);
            dxm = ldg(A.dx1, ob); dym = ldg(A.dy1, ob); dzm = ldg(A.dz1, ob);
            exm = ldg(A.ex1, ob); eym = ldg(A.ey1, ob); ezm = ldg(A.ez1, ob);
            id = A.ids[ob >> 2];
        }
        // Halo H^{n-1/2}: row j0-1 (Hx, Hz) and column i0-1 (Hy, Hz).
        float hrx = 0.f, hrz = 0.f, hcy = 0.f, hcz = 0.f;
        if (row_halo) {
            const uint32_t oh = (uint32_t)k * szb + rowb;
            hrx = ldg(A.hxo, oh); hrz = ldg(A.hzo, oh);
        }
        if (col_halo) {
            const uint32_t oh = (uint32_t)k * szb + colhb;
            hcy = ldg(A.hyo, oh); hcz = ldg(A.hzo, oh);
        }

        load_plane(nxt, k + 1);
        __syncthreads();  // planes k (cur) and k+1 (nxt) staged

        const float exk = sE[cur][0][ty + 1][tx + 1], eyk = sE[cur][1][ty + 1][tx + 1], ezk = sE[cur][2][ty + 1][tx + 1];
        float hx = hxo, hy = hyo, hz = hzo;
        if (own) {
            const float exk1 = sE[nxt][0][ty + 1][tx + 1], eyk1 = sE[nxt][1][ty + 1][tx + 1];
            if (i >= 1 && i < nx - 1 && j < ny - 1 && k < nz - 1)
                hx = hx - chy * (sE[cur][2][ty + 2][tx + 1] - ezk) + chy * (eyk1 - eyk);
            if (i < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1)
                hy = hy - chz * (exk1 - exk) + chx * (sE[cur][2][ty + 1][tx + 2] - ezk);
            if (i < nx - 1 && j < ny - 1 && k >= 1 && k < nz - 1)
                hz = hz - chx * (sE[cur][1][ty + 1][tx + 2] - eyk) + chy * (sE[cur][0][ty + 2][tx + 1] - exk);
        }
        sH[0][ty + 1][tx + 1] = hx;
        sH[1][ty + 1][tx + 1] = hy;
        sH[2][ty + 1][tx + 1] = hz;
        if (ty == 0) {
            const int jh = j0 - 1;
            if (row_halo) {
                const float ezh = sE[cur][2][0][tx + 1], eyh = sE[cur][1][0][tx + 1], exh = sE[cur][0][0][tx + 1];
                if (i >= 1 && i < nx - 1 && jh < ny - 1 && k < nz - 1)
                    hrx = hrx - chy * (sE[cur][2][1][tx + 1] - ezh) + chy * (sE[nxt][1][0][tx + 1] - eyh);
                if (i < nx - 1 && jh < ny - 1 && k >= 1 && k < nz - 1)
                    hrz = hrz - chx * (sE[cur][1][0][tx + 2] - eyh) + chy * (sE[cur][0][1][tx + 1] - exh);
            }
            sH[0][0][tx + 1] = hrx;
            sH[2][0][tx + 1] = hrz;
        }
        if (tx == 0) {
            const int ih = i0 - 1;
            if (col_halo) {
                const float exh = sE[cur][0][ty + 1][0], eyh = sE[cur][1][ty + 1][0], ezh = sE[cur][2][ty + 1][0];
                if (ih < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1)
                    hcy = hcy - chz * (sE[nxt][0][ty + 1][0] - exh) + chx * (sE[cur][2][ty + 1][1] - ezh);
                if (ih < nx - 1 && j < ny - 1 && k >= 1 && k < nz - 1)
                    hcz = hcz - chx * (sE[cur][1][ty + 1][1] - eyh) + chy * (sE[cur][0][ty + 2][0] - exh);
            }
            sH[1][ty + 1][0] = hcy;
            sH[2][ty + 1][0] = hcz;
        }
        if (own) {
            stg(A.hxn, ob, hx); stg(A.hyn, ob, hy); stg(A.hzn, ob, hz);
        }
        __syncthreads();  // H^{n+1/2} of plane k and its halo staged

        if (own) {
            const bool rdx = (i < nx - 1 && j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1);
            const bool rdy = (i >= 1 && i < nx - 1 && j < ny - 1 && k >= 1 && k < nz - 1);
            const bool rdz = (i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1);
            const bool reyz = (i >= 1 && k >= 1);
            float dxp = 0.f, dyp = 0.f, dzp = 0.f;
            if (rdx) dxp = dxn + cdy * (hz - sH[2][ty][tx + 1]) - cdz * (hy - hym);
            if (rdy) dyp = dyn + cdz * (hx - hxm) - cdx * (hz - sH[2][ty + 1][tx]);
            if (rdz) {
                dzp = dzn + cdx * (hy - sH[1][ty + 1][tx]) - cdy * (hx - sH[0][ty][tx + 1]);
                if (id & kSourceBit) dzp = dzp - A.jz;
            }
            // Cells the D update skips keep the target slot's value.
            if (reyz && !rdy) dyp = ldg(A.dyT, ob);
            if (reyz && !rdz) dzp = ldg(A.dzT, ob);
            if (rdx) stg(A.dxT, ob, dxp);
            if (rdy) stg(A.dyT, ob, dyp);
            if (rdz) stg(A.dzT, ob, dzp);
            const int m = id & kMaterialMask;
            const float ca = lA[m], cc = lC[m], cd = lD[m];
            const double g1 = lG1[m], g2 = lG2[m], g3 = lG3[m];
            if (rdx) stg(A.exT, ob, e_update(cc, cd, ca, g1, g2, g3, exk, exm, dxp, dxn, dxm));
            if (reyz) {
                stg(A.eyT, ob, e_update(cc, cd, ca, g1, g2, g3, eyk, eym, dyp, dyn, dym));
                stg(A.ezT, ob, e_update(cc, cd, ca, g1, g2, g3, ezk, ezm, dzp, dzn, dzm));
            }
        }
        hxm = hx;
        hym = hy;
        cur = nxt;
    }
}

__device__ inline float mur1(float dt, float h, float mu0, float eps0, float rpf, float rpn, float tn, float sn,
                             float sf)
{
    return 1 / (dt + h * sqrtf(mu0 * eps0 * rpf)) *
           ((dt - h * sqrtf(mu0 * eps0 * rpn)) * tn + (dt + h * sqrtf(mu0 * eps0 * rpn)) * sn -
            (dt - h * sqrtf(mu0 * eps0 * rpf)) * sf);
}

constexpr int kMurBlock = 256;

__global__ __launch_bounds__(kMurBlock) void mur_faces(const MurArgs M)
{
    const int t = blockIdx.x * kMurBlock + threadIdx.x;
    if (t >= M.seg_start[12]) return;
    int s = 0;
#pragma unroll
    for (int q = 1; q < 12; q++) s += (t >= M.seg_start[q]) ? 1 : 0;
    const int q = t - M.seg_start[s];
    const int nx = M.nx, ny = M.ny, nz = M.nz;
    const long long sy = M.pitch, sz = M.plane;
    const float dt = M.dt, m0 = M.mu0, e0 = M.eps0;
    auto at = [&](int i, int j, int k) { return (long long)k * sz + (long long)j * sy + i; };
    auto rpy = [&](const float *p, int i, int jj, int k) { return p[(long long)k * 2 * nx + (long long)jj * nx + i]; };
    auto rpz = [&](const float *p, int i, int j, int kk) { return p[(long long)kk * ny * nx + (long long)j * nx + i]; };
    auto rpx = [&](const float *p, int ii, int j, int k) { return p[(long long)k * ny * 2 + (long long)j * 2 + ii]; };
    switch (s) {
    case 0: case 1: {  // Ex on y faces: i in [0,nx-1), k in [1,nz-1)
        const int i = q % (nx - 1), k = 1 + q / (nx - 1);
        if (s == 0) {
            const long long o = at(i, 0, k);
            M.exT[o] = mur1(dt, M.hy, m0, e0, rpy(M.rpy0, i, 0, k), rpy(M.rpy0, i, 1, k), M.exT[o + sy], M.exS[o + sy], M.exS[o]);
        } else {
            const long long o = at(i, ny - 1, k);
            M.exT[o] = mur1(dt, M.hy, m0, e0, rpy(M.rpyEnd, i, 1, k), rpy(M.rpyEnd, i, 0, k), M.exT[o - sy], M.exS[o - sy], M.exS[o]);
        }
        break;
    }
    case 2: case 3: {  // Ex on z faces: i in [0,nx-1), j in [1,ny-1)
        const int i = q % (nx - 1), j = 1 + q / (nx - 1);
        if (s == 2) {
            const long long o = at(i, j, 0);
            M.exT[o] = mur1(dt, M.hz, m0, e0, rpz(M.rpz0, i, j, 0), rpz(M.rpz0, i, j, 1), M.exT[o + sz], M.exS[o + sz], M.exS[o]);
        } else {
            const long long o = at(i, j, nz - 1);
            M.exT[o] = mur1(dt, M.hz, m0, e0, rpz(M.rpzEnd, i, j, 1), rpz(M.rpzEnd, i, j, 0), M.exT[o - sz], M.exS[o - sz], M.exS[o]);
        }
        break;
    }
    case 4: case 5: {  // Ey on x faces: j in [0,ny-1), k in [1,nz-1)
        const int j = q % (ny - 1), k = 1 + q / (ny - 1);
        if (s == 4) {
            const long long o = at(0, j, k);
            M.eyT[o] = mur1(dt, M.hx, m0, e0, rpx(M.rpx0, 0, j, k), rpx(M.rpx0, 1, j, k), M.eyT[o + 1], M.eyS[o + 1], M.eyS[o]);
        } else {  // the x-high Ey face reads the source neighbour twice (fdtd_calculations.cu:323-325)
            const long long o = at(nx - 1, j, k);
            M.eyT[o] = mur1(dt, M.hx, m0, e0, rpx(M.rpxEnd, 1, j, k), rpx(M.rpxEnd, 0, j, k), M.eyS[o - 1], M.eyS[o - 1], M.eyS[o]);
        }
        break;
    }
    case 6: case 7: {  // Ey on z faces: i in [1,nx-1), j in [0,ny-1)
        const int i = 1 + q % (nx - 2), j = q / (nx - 2);
        if (s == 6) {
            const long long o = at(i, j, 0);
            M.eyT[o] = mur1(dt, M.hz, m0, e0, rpz(M.rpz0, i, j, 0), rpz(M.rpz0, i, j, 1), M.eyT[o + sz], M.eyS[o + sz], M.eyS[o]);
        } else {
            const long long o = at(i, j, nz - 1);
            M.eyT[o] = mur1(dt, M.hz, m0, e0, rpz(M.rpzEnd, i, j, 1), rpz(M.rpzEnd, i, j, 0), M.eyT[o - sz], M.eyS[o - sz], M.eyS[o]);
        }
        break;
    }
    case 8: case 9: {  // Ez on x faces: j in [1,ny-1), k in [0,nz-1)
        const int j = 1 + q % (ny - 2), k = q / (ny - 2);
        if (s == 8) {
            const long long o = at(0, j, k);
            M.ezT[o] = mur1(dt, M.hx, m0, e0, rpx(M.rpx0, 0, j, k), rpx(M.rpx0, 1, j, k), M.ezT[o + 1], M.ezS[o + 1], M.ezS[o]);
        } else {
            const long long o = at(nx - 1, j, k);
            M.ezT[o] = mur1(dt, M.hx, m0, e0, rpx(M.rpxEnd, 1, j, k), rpx(M.rpxEnd, 0, j, k), M.ezT[o - 1], M.ezS[o - 1], M.ezS[o]);
        }
        break;
    }
    default: {  // Ez on y faces: i in [1,nx-1), k in [0,nz-1)
        const int i = 1 + q % (nx - 2), k = q / (nx - 2);
        if (s == 10) {
            const long long o = at(i, 0, k);
            M.ezT[o] = mur1(dt, M.hy, m0, e0, rpy(M.rpy0, i, 0, k), rpy(M.rpy0, i, 1, k), M.ezT[o + sy], M.ezS[o + sy], M.ezS[o]);
        } else {
            const long long o = at(i, ny - 1, k);
            M.ezT[o] = mur1(dt, M.hy, m0, e0, rpy(M.rpyEnd, i, 1, k), rpy(M.rpyEnd, i, 0, k), M.ezT[o - sy], M.ezS[o - sy], M.ezS[o]);
        }
        break;
    }
    }
}

__global__ void probe_cells(const ProbeArgs P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.nprobe) return;
    const int *q = P.probes + 4 * p;
    const long long o = (long long)q[3] * P.plane + (long long)q[2] * P.pitch + q[1];
    P.out[(long long)P.row * P.nprobe + p] = (double)P.field[q[0]][o];
}

}  // namespace

void mur_segments(int nx, int ny, int nz, int *seg)
{
    const int cnt[12] = {
        (nx - 1) * (nz - 2), (nx - 1) * (nz - 2), (nx - 1) * (ny - 2), (nx - 1) * (ny - 2),
        (ny - 1) * (nz - 2), (ny - 1) * (nz - 2), (nx - 2) * (ny - 1), (nx - 2) * (ny - 1),
        (ny - 2) * (nz - 1), (ny - 2) * (nz - 1), (nx - 2) * (nz - 1), (nx - 2) * (nz - 1),
    };
    seg[0] = 0;
    for (int q = 0; q < 12; q++) seg[q + 1] = seg[q] + std::max(cnt[q], 0);
}

hipError_t launch_mur(const MurArgs &a, hipStream_t s)
{
    const int total = a.seg_start[12];
    if (total <= 0) return hipSuccess;
    mur_faces<<<(total + kMurBlock - 1) / kMurBlock, kMurBlock, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_probes(const ProbeArgs &a, hipStream_t s)
{
    if (a.nprobe <= 0) return hipSuccess;
    probe_cells<<<(a.nprobe + 63) / 64, 64, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_fused_step(const StepArgs &a, int ty, hipStream_t s)
{
    const unsigned nblk = (unsigned)a.tiles_x * a.tiles_y * a.tiles_z;
    switch (ty) {
    case 1: fused_step<1><<<nblk, dim3(TX, 1), 0, s>>>(a); break;
    case 2: fused_step<2><<<nblk, dim3(TX, 2), 0, s>>>(a); break;
    case 4: fused_step<4><<<nblk, dim3(TX, 4), 0, s>>>(a); break;
    case 8: fused_step<8><<<nblk, dim3(TX, 8), 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
