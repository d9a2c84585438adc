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
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "fdtd_internal.h"

namespace {

constexpr int TX = 64;
constexpr int kMurK = kFusedMurMaxChunk;  // z planes of staged Mur coefficients per tile

__device__ inline int xcd_logical(int b, int nblk, int mode, int unit)
{
    // Workgroups are dispatched to the 8 XCDs round-robin by id (b % 8).
    if (mode == 1) {
        // Each XCD takes one contiguous run of logical tiles.
        const int q = nblk / 8, rem = nblk % 8;
        const int x = b % 8, s = b / 8;
        return x * q + min(x, rem) + s;
    }
    if (mode >= 2) {
        // Groups of 8 units: XCD x takes unit x of each group.  Mode 2: a
        // unit is one z-chunk layer of tiles, so the 8 XCDs work on 8
        // adjacent layers while the tiles of a layer (which share halo rows
        // and columns) share one L2.  Mode 3: a unit is one row of tiles
        // along x, so the tiles that share x-halo columns share an L2 and
        // the 8 XCDs work on adjacent rows of one layer.  Mode 4: a unit is
        // an eighth of a layer's rows (a strip), so the XCDs split each
        // layer and x and y neighbours share an L2 inside a strip.  A
        // partial last group keeps the identity.
        const int G = 8 * unit;
        const int full = (nblk / G) * G;
        if (b >= full) return b;
        const int g = b / G, r = b % G;
        return g * G + (r % 8) * unit + r / 8;
    }
    return b;
}

// Tiles per unit of xcd_logical's modes 2-4.
__device__ inline int xcd_unit(int mode, int tiles_x, int tiles_y)
{
    return mode == 3 ? tiles_x : mode == 4 ? (tiles_y + 7) / 8 * tiles_x : tiles_x * tiles_y;
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

// Element access at a 32-bit byte offset from a scalar base, for the three
// storage types.  fp32 values pass through; fp16/bf16 convert to and from
// fp32 with round-to-nearest-even.  The _nt forms are the streaming
// (nontemporal) variants for data a launch touches once: the ring levels n
// and n-1 of D, level n-1 of E and every store.
template <typename T> struct Elem;

template <> struct Elem<float> {
    static __device__ __forceinline__ float rnd(float v) { return v; }
    static __device__ __forceinline__ float ld(const float *p, uint32_t b) { return *(const float *)((const char *)p + b); }
    static __device__ __forceinline__ float ld_nt(const float *p, uint32_t b)
    {
        return __builtin_nontemporal_load((const float *)((const char *)p + b));
    }
    static __device__ __forceinline__ void st(float *p, uint32_t b, float v) { *(float *)((char *)p + b) = v; }
    static __device__ __forceinline__ void st_nt(float *p, uint32_t b, float v)
    {
        __builtin_nontemporal_store(v, (float *)((char *)p + b));
    }
};

template <> struct Elem<__half> {
    static __device__ __forceinline__ float cvt(unsigned short u) { return __half2float(__ushort_as_half(u)); }
    static __device__ __forceinline__ unsigned short enc(float v) { return __half_as_ushort(__float2half_rn(v)); }
    static __device__ __forceinline__ float rnd(float v) { return cvt(enc(v)); }
    static __device__ __forceinline__ float ld(const __half *p, uint32_t b)
    {
        return cvt(*(const unsigned short *)((const char *)p + b));
    }
    static __device__ __forceinline__ float ld_nt(const __half *p, uint32_t b)
    {
        return cvt(__builtin_nontemporal_load((const unsigned short *)((const char *)p + b)));
    }
    static __device__ __forceinline__ void st(__half *p, uint32_t b, float v) { *(unsigned short *)((char *)p + b) = enc(v); }
    static __device__ __forceinline__ void st_nt(__half *p, uint32_t b, float v)
    {
        __builtin_nontemporal_store(enc(v), (unsigned short *)((char *)p + b));
    }
};

template <> struct Elem<__hip_bfloat16> {
    static __device__ __forceinline__ float cvt(unsigned short u) { return __uint_as_float((uint32_t)u << 16); }
    static __device__ __forceinline__ unsigned short enc(float v)
    {
        // round to nearest even; NaN stays NaN
        const uint32_t x = __float_as_uint(v);
        if ((x & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((x >> 16) | 0x40u);
        return (unsigned short)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
    }
    static __device__ __forceinline__ float rnd(float v) { return cvt(enc(v)); }
    static __device__ __forceinline__ float ld(const __hip_bfloat16 *p, uint32_t b)
    {
        return cvt(*(const unsigned short *)((const char *)p + b));
    }
    static __device__ __forceinline__ float ld_nt(const __hip_bfloat16 *p, uint32_t b)
    {
        return cvt(__builtin_nontemporal_load((const unsigned short *)((const char *)p + b)));
    }
    static __device__ __forceinline__ void st(__hip_bfloat16 *p, uint32_t b, float v)
    {
        *(unsigned short *)((char *)p + b) = enc(v);
    }
    static __device__ __forceinline__ void st_nt(__hip_bfloat16 *p, uint32_t b, float v)
    {
        __builtin_nontemporal_store(enc(v), (unsigned short *)((char *)p + b));
    }
};

// E^{n+1} in scaled units (16-bit storage): fp32 throughout.
__device__ inline float e_update_scaled(float k1, float k2, float k3, float k5, float e0, float e1, float dn1,
                                        float dn, float dm1)
{
    const float d1 = dn1 - dn;
    const float d2 = d1 - (dn - dm1);
    return k1 * e0 - k2 * e1 + k3 * d1 + k5 * d2;
}

// First-order Mur update of one face cell (fdtd_calculations.cu:237-427):
// rpf/rpn are the face and neighbour permittivities, tn the neighbour at
// n+1, sn and sf the neighbour and the face at n.
__device__ inline float mur1(float dt, float h, float mu0, float eps0, float rpf, float rpn, float tn, float sn,
                             float sf)
{
    return 1 / (dt + h * sqrtf(mu0 * eps0 * rpf)) *
           ((dt - h * sqrtf(mu0 * eps0 * rpn)) * tn + (dt + h * sqrtf(mu0 * eps0 * rpn)) * sn -
            (dt - h * sqrtf(mu0 * eps0 * rpf)) * sf);
}

// All per-lane addresses are a 64-bit array base (uniform, in SGPRs) plus a
// 32-bit byte offset shared by every array indexed by the same cell, so the
// loads use the scalar-base form with one offset register.  Offsets are
// measured from `guard` elements before each array's origin: every array
// (and the id array) has at least `guard` addressable elements before and
// after it, so the halo loads at i-1, i+64, j-1, j+TY, k-1 and k+1 need no
// bounds checks (their values are discarded outside the update ranges).
//
// Roles inside a block (wave-uniform): wave 0 also loads and updates the
// -y halo row, wave TY-1 loads the +y E row, and wave 1 gathers the -x/+x
// E columns and updates the -x halo H column (lane l: side l/32, row l%32).
// Each plane's global loads are issued one plane ahead of their use.
// The per-material coefficient table lives in dynamic LDS: exact form g1 g2
// g3 [nmat] doubles then a c d [nmat] floats; scaled form k1 k2 k3 k5 floats.
__device__ __forceinline__ double *lut_lds()
{
    extern __shared__ double sLut[];
    return sLut;
}

template <bool EXACT>
__device__ __forceinline__ void stage_lut(const MatCoef *lut, int nmat, int tid, int nthr)
{
    double *s = lut_lds();
    double *lG1 = s, *lG2 = s + nmat, *lG3 = s + 2 * nmat;
    float *lA = (float *)(s + 3 * nmat), *lC = lA + nmat, *lD = lA + 2 * nmat;
    float *lK1 = (float *)s, *lK2 = lK1 + nmat, *lK3 = lK1 + 2 * nmat, *lK5 = lK1 + 3 * nmat;
    for (int m = tid; m < nmat; m += nthr) {
        const MatCoef mc = lut[m];
        if constexpr (EXACT) {
            lA[m] = mc.a; lC[m] = mc.c; lD[m] = mc.d;
            lG1[m] = mc.g1; lG2[m] = mc.g2; lG3[m] = mc.g3;
        } else {
            lK1[m] = mc.k1; lK2[m] = mc.k2; lK3[m] = mc.k3; lK5[m] = mc.k5;
        }
    }
}

// One TXW x TY x kchunk tile (column tile (bx, by), z-chunk bz) of one step,
// TXW = 64 (one wave per row) or 128 (two waves per row, which halves the
// x-halo column reads per cell).  The coefficient table must be staged; the
// first LDS read of it follows a barrier inside.  The tile's own LDS writes
// start at once, so a caller that runs several tiles per block puts a
// barrier between them.
template <int TY, typename T, bool EXACT, int TXW = 64, bool ZS = false>
__device__ __forceinline__ void fused_tile(const StepArgs &A, int bx, int by, int bz)
{
    static_assert(TXW == 64 || TXW == 128, "tile width");
    static_assert(!ZS || EXACT, "zero-term rows belong to the exact form");
    static_assert(TY >= 2 && TY <= 8, "tile rows");
    static_assert(!EXACT || std::is_same<T, float>::value, "the exact form stores fp32");
    // EXACT reproduces the reference's arithmetic (double Debye sums); the
    // other form runs scaled units in fp32 (16-bit storage, fast fp32).
    constexpr bool exact = EXACT;
    constexpr uint32_t ES = sizeof(T);
    using X = Elem<T>;
    __shared__ float sE[2][3][TY + 2][TXW + 2];
    __shared__ float sH[3][TY + 1][TXW + 1];
    __shared__ float sJunk[2][TXW];

    // every wave covers one tile row (64 lanes along x), so the row index is
    // wave-uniform: as a scalar it keeps the row tests and the wave roles
    // (halo rows, column gather) out of the per-lane masks
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int i0 = bx * TXW, j0 = by * TY;
    const int k0 = bz * A.kchunk;
    const int k1 = min(k0 + A.kchunk, nz);
    const int i = i0 + tx, j = j0 + ty;
    const bool own = (i < nx) && (j < ny);

    const int nmat = A.nmat;
    double *sLut = lut_lds();
    double *lG1 = sLut, *lG2 = sLut + nmat, *lG3 = sLut + 2 * nmat;
    float *lA = (float *)(sLut + 3 * nmat), *lC = lA + nmat, *lD = lA + 2 * nmat;
    float *lK1 = (float *)sLut, *lK2 = lK1 + nmat, *lK3 = lK1 + 2 * nmat, *lK5 = lK1 + 3 * nmat;

    // Origin-shifted bases (guard elements before each array's origin).
    const uint32_t gb = A.guard * ES;
    auto sh = [&](const void *p) { return (const T *)((const char *)p - gb); };
    auto shw = [&](void *p) { return (T *)((char *)p - gb); };
    const T *hxo = sh(A.hxo), *hyo = sh(A.hyo), *hzo = sh(A.hzo);
    const T *ex0 = sh(A.ex0), *ey0 = sh(A.ey0), *ez0 = sh(A.ez0);
    const T *ex1 = sh(A.ex1), *ey1 = sh(A.ey1), *ez1 = sh(A.ez1);
    const T *dx0 = sh(A.dx0), *dy0 = sh(A.dy0), *dz0 = sh(A.dz0);
    const T *dx1 = sh(A.dx1), *dy1 = sh(A.dy1), *dz1 = sh(A.dz1);
    T *hxn = shw(A.hxn), *hyn = shw(A.hyn), *hzn = shw(A.hzn);
    T *exT = shw(A.exT), *eyT = shw(A.eyT), *ezT = shw(A.ezT);
    T *dxT = shw(A.dxT), *dyT = shw(A.dyT), *dzT = shw(A.dzT);
    const uint8_t *ids = A.ids - A.guard;

    const uint32_t syb = (uint32_t)A.pitch * ES, szb = (uint32_t)A.plane * ES;
    // Lanes past the row pitch (the last x tile of a row whose pitch is not a
    // multiple of 64) read column pitch-1 instead: a line the lanes below
    // them read anyway.  Their values are never used.  (A branch around the
    // loads instead would make the compiler wait for each load inside it.)
    const int il = min(i, A.pitch - 1);
    // plane-relative byte offsets (wrap-around arithmetic handles j0-1 = -1 etc.)
    const uint32_t c_own = gb + (uint32_t)j * syb + (uint32_t)il * ES;
    const uint32_t c_rlo = gb + (uint32_t)(j0 - 1) * syb + (uint32_t)il * ES;
    const uint32_t c_rhi = gb + (uint32_t)(j0 + TY) * syb + (uint32_t)il * ES;
    const int side = tx >> 5;
    const int cr = min(tx & 31, TY + 1);             // LDS row of the gathered column cell
    const int cl = side ? TXW + 1 : 0;                // LDS column of the gathered cell
    const int ci = side ? i0 + TXW : i0 - 1;
    const uint32_t c_col = gb + (uint32_t)(j0 - 1 + cr) * syb + (uint32_t)ci * ES;
    // the column gather is the first wave of row 1 (wave-uniform for both widths)
    const bool w_rlo = (ty == 0), w_rhi = (ty == TY - 1), w_col = (ty == 1) && (tx < 64);

    const float chx = A.chx, chy = A.chy, chz = A.chz;
    const float cdx = A.cdx, cdy = A.cdy, cdz = A.cdz;

    // E^n staging registers of one plane: own cell, -y row (wave 0), +y row
    // (wave TY-1, Ex and Ez), gathered columns (wave 1).
    float pe[3], pr[3], pq[2], pc[3];
    auto fetch_e = [&](int k) {
        const uint32_t kb = (uint32_t)k * szb;
        pe[0] = X::ld(ex0, kb + c_own); pe[1] = X::ld(ey0, kb + c_own); pe[2] = X::ld(ez0, kb + c_own);
        if (w_rlo) { pr[0] = X::ld(ex0, kb + c_rlo); pr[1] = X::ld(ey0, kb + c_rlo); pr[2] = X::ld(ez0, kb + c_rlo); }
        if (w_rhi) { pq[0] = X::ld(ex0, kb + c_rhi); pq[1] = X::ld(ez0, kb + c_rhi); }
        if (w_col) { pc[0] = X::ld(ex0, kb + c_col); pc[1] = X::ld(ey0, kb + c_col); pc[2] = X::ld(ez0, kb + c_col); }
    };
    auto stage_e = [&](int buf) {
        sE[buf][0][ty + 1][tx + 1] = pe[0]; sE[buf][1][ty + 1][tx + 1] = pe[1]; sE[buf][2][ty + 1][tx + 1] = pe[2];
        if (w_rlo) { sE[buf][0][0][tx + 1] = pr[0]; sE[buf][1][0][tx + 1] = pr[1]; sE[buf][2][0][tx + 1] = pr[2]; }
        if (w_rhi) { sE[buf][0][TY + 1][tx + 1] = pq[0]; sE[buf][2][TY + 1][tx + 1] = pq[1]; }
        if (w_col) { sE[buf][0][cr][cl] = pc[0]; sE[buf][1][cr][cl] = pc[1]; sE[buf][2][cr][cl] = pc[2]; }
    };

    // Own-cell operands of one plane.  zs: E^{n-1} and D^{n-1} were not
    // loaded (zero-term row, below) and read as +0.
    struct Own {
        float hxo, hyo, hzo, dxn, dyn, dzn, dxm, dym, dzm, exm, eym, ezm, hh0, hh1;
        uint32_t id;
        bool zs;
    };
    // ids load cached: a wave's 64 ids are half of a 128-byte line whose other half the x-neighbour
    // tile (on the same XCD in the row block order) reads
    auto fetch_id = [&](int k) -> uint32_t { return ids[((uint32_t)k * szb + c_own) / ES]; };
    // load_id: the plane's ids come with its operands (else the caller loaded them ahead)
    auto fetch_own = [&](int k, Own &o, bool zs, bool load_id) {
        const uint32_t ob = (uint32_t)k * szb + c_own;
        o.hxo = X::ld(hxo, ob); o.hyo = X::ld(hyo, ob); o.hzo = X::ld(hzo, ob);
        o.dxn = X::ld_nt(dx0, ob); o.dyn = X::ld_nt(dy0, ob); o.dzn = X::ld_nt(dz0, ob);
        // Loaded into locals and assigned to the struct unconditionally: with the
        // members assigned on both branches, the compiler merged the two paths'
        // last stores into one through a selected member address and kept part
        // of the struct in scratch memory, written back to HBM every plane.
        float dxm = 0.0f, dym = 0.0f, dzm = 0.0f, exm = 0.0f, eym = 0.0f, ezm = 0.0f;
        if (!(ZS && zs)) {
            dxm = X::ld_nt(dx1, ob); dym = X::ld_nt(dy1, ob); dzm = X::ld_nt(dz1, ob);
            exm = X::ld_nt(ex1, ob); eym = X::ld_nt(ey1, ob); ezm = X::ld_nt(ez1, ob);
        }
        o.dxm = dxm; o.dym = dym; o.dzm = dzm;
        o.exm = exm; o.eym = eym; o.ezm = ezm;
        if (load_id) o.id = ids[ob / ES];
        o.zs = ZS && zs;
        // halo H^{n-1/2}: row (i, j0-1): Hx, Hz; column (i0-1 | i0+TXW, row cr): Hy, Hz
        o.hh0 = 0.f; o.hh1 = 0.f;
        if (w_rlo) { const uint32_t oh = (uint32_t)k * szb + c_rlo; o.hh0 = X::ld(hxo, oh); o.hh1 = X::ld(hzo, oh); }
        if (w_col) { const uint32_t oh = (uint32_t)k * szb + c_col; o.hh0 = X::ld(hyo, oh); o.hh1 = X::ld(hzo, oh); }
    };
    // Zero-term rows (ZS, exact form): a material with tau = +0 has a = +0 and
    // g3 = +0, so its Debye update multiplies E^{n-1} and D^{n-1} by +0.  When
    // every own cell of a half wave (32 cells, one 128-byte line of each fp32
    // array) has such a material, its six E^{n-1}/D^{n-1} loads are skipped
    // (uniform per half: the other half's loads fetch only their own lines).
    // The ids move two planes ahead: the id load of plane k+2 is issued first
    // in iteration k and decided at its end, so plane k+2's loads go out at
    // once in iteration k+1.  The first z-chunk decides from its third plane
    // on (no barrier guards the coefficient table earlier), the others from
    // their second.
    const bool zskip = ZS && A.zskip;
    auto zero_row = [&](uint32_t id) -> bool {
        const int m = id & kMaterialMask;
        const bool z = __float_as_uint(lA[m]) == 0u && __double_as_longlong(lG3[m]) == 0ll;   // both +0
        const uint64_t b = __ballot(!own || z);
        return ((b >> (tx & 32)) & 0xffffffffull) == 0xffffffffull;
    };

    // Mur x faces: the tile's coefficients (planes k0..k1-1, rows j0..j0+TY-1,
    // both faces) and, on the k = 0 chunk, the never-updated Ez^{n+1}
    // neighbours at k = 0 go to LDS up front, so the loop issues no global
    // load beyond the streams.
    __shared__ float4 sMurC[2][kMurK][TY];
    __shared__ float sMurT[2][TY];
    const bool mur_lo = A.mur_x && i0 == 0;
    const bool mur_hi = A.mur_x == 1 && i0 <= nx - 1 && nx - 1 < i0 + TXW;
    if (mur_lo || mur_hi) {
        const int tid = ty * TXW + tx, nk = k1 - k0;
        for (int q = tid; q < 2 * nk * TY; q += TXW * TY) {
            const int side = q / (nk * TY), r = q - side * nk * TY, kk = r / TY, row = r - kk * TY;
            const long long at = (long long)(k0 + kk) * ny + min(j0 + row, ny - 1);
            sMurC[side][kk][row] = side ? A.murx1[at] : A.murx0[at];
        }
        if (k0 == 0 && tid < 2 * TY) {
            const int side = tid / TY, row = tid - side * TY;
            const int ii = side ? nx - 2 : 1, jj = min(j0 + row, ny - 1);
            sMurT[side][row] = X::ld(ezT, gb + (uint32_t)jj * syb + (uint32_t)ii * ES);
        }
    }

    float hxm = 0.0f, hym = 0.0f;  // own-column Hx, Hy of plane k-1 (time level n+1/2)
    int cur = 0;
    uint32_t idq = 0;     // ZS: ids of the plane after next, loaded ahead
    bool zsn = false;     // ZS: the next plane skips its E^{n-1}, D^{n-1} loads
    if constexpr (ZS) {
        if (k0 + 1 < k1) idq = fetch_id(k0 + 1);
    }
    if (k0 > 0) {
        fetch_e(k0 - 1);
        stage_e(0);
        fetch_e(k0);
        stage_e(1);
        __syncthreads();
        if constexpr (ZS) zsn = zskip && k0 + 1 < k1 && zero_row(idq);   // table staged, idq in
        const int k = k0 - 1;
        const uint32_t ob = (uint32_t)k * szb + c_own;
        const float hx0 = X::ld(hxo, ob), hy0 = X::ld(hyo, ob);
        const float exk = sE[0][0][ty + 1][tx + 1], eyk = sE[0][1][ty + 1][tx + 1], ezk = sE[0][2][ty + 1][tx + 1];
        const float exk1 = sE[1][0][ty + 1][tx + 1], eyk1 = sE[1][1][ty + 1][tx + 1];
        const float ux = hx0 - chy * (sE[0][2][ty + 2][tx + 1] - ezk) + chy * (eyk1 - eyk);
        const float uy = hy0 - chz * (exk1 - exk) + chx * (sE[0][2][ty + 1][tx + 2] - ezk);
        hxm = (i >= 1 && i < nx - 1 && j < ny - 1 && k < nz - 1) ? ux : hx0;
        hym = (i < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1) ? uy : hy0;
        cur = 1;
        __syncthreads();
    } else {
        fetch_e(k0);
        stage_e(0);
    }
    Own cu, nx_;
    fetch_own(k0, cu, false, true);
    fetch_e(k0 + 1);

    // Range masks that do not depend on k.
    const bool hx_ij = (i >= 1 && i < nx - 1 && j < ny - 1);
    const bool hy_ij = (i < nx - 1 && j >= 1 && j < ny - 1);
    const bool hz_ij = (i < nx - 1 && j < ny - 1);
    const bool dx_ij = (i < nx - 1 && j >= 1 && j < ny - 1);
    const bool dy_ij = (i >= 1 && i < nx - 1 && j < ny - 1);
    const bool dz_ij = (i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1);
    const bool e_i = (i >= 1);
    // halo row (i, j0-1) and halo column (i0-1, j0-1+cr) masks
    const int jh = j0 - 1;
    const bool rx_ij = (jh >= 0) && (i >= 1 && i < nx - 1 && jh < ny - 1);
    const bool rz_ij = (jh >= 0) && (i < nx - 1 && jh < ny - 1);
    const int ccr = min(max(cr, 1), TY);              // clamped LDS row for the column compute
    const int jc = j0 - 1 + ccr;
    const bool col_ok = (side == 0) && (cr >= 1) && (cr <= TY) && (i0 >= 1);
    const bool cy_ij = col_ok && (i0 - 1 < nx - 1) && (jc >= 1 && jc < ny - 1);
    const bool cz_ij = col_ok && (i0 - 1 < nx - 1) && (jc < ny - 1);
    float *col_dst_y = col_ok ? &sH[1][ccr][0] : &sJunk[0][tx];
    float *col_dst_z = col_ok ? &sH[2][ccr][0] : &sJunk[1][tx];

    const float jzv = A.jz;
    for (int k = k0; k < k1; ++k) {
        const int nxt = cur ^ 1;
        stage_e(nxt);                 // E^n of plane k+1
        const bool more = (k + 1 < k1);
        if (more) {
            if constexpr (ZS) {
                nx_.id = idq;                                 // plane k+1
                if (k + 2 < k1) idq = fetch_id(k + 2);
                fetch_own(k + 1, nx_, zsn, false);
            } else {
                fetch_own(k + 1, nx_, false, true);
            }
            fetch_e(k + 2);
        }
        __syncthreads();  // planes k (cur) and k+1 (nxt) staged

        const bool kh = (k < nz - 1), kz = (k >= 1 && k < nz - 1);
        // own H^{n+1/2}
        const float exk = sE[cur][0][ty + 1][tx + 1], eyk = sE[cur][1][ty + 1][tx + 1], ezk = sE[cur][2][ty + 1][tx + 1];
        const float exk1 = sE[nxt][0][ty + 1][tx + 1], eyk1 = sE[nxt][1][ty + 1][tx + 1];
        const float ux = cu.hxo - chy * (sE[cur][2][ty + 2][tx + 1] - ezk) + chy * (eyk1 - eyk);
        const float uy = cu.hyo - chz * (exk1 - exk) + chx * (sE[cur][2][ty + 1][tx + 2] - ezk);
        const float uz = cu.hzo - chx * (sE[cur][1][ty + 1][tx + 2] - eyk) + chy * (sE[cur][0][ty + 2][tx + 1] - exk);
        const float hx = (hx_ij && kh) ? ux : cu.hxo;
        const float hy = (hy_ij && kh) ? uy : cu.hyo;
        const float hz = (hz_ij && kz) ? uz : cu.hzo;
        sH[0][ty + 1][tx + 1] = hx;
        sH[1][ty + 1][tx + 1] = hy;
        sH[2][ty + 1][tx + 1] = hz;
        if (w_rlo) {  // (i, j0-1): Hx, Hz
            const float ezh = sE[cur][2][0][tx + 1], eyh = sE[cur][1][0][tx + 1], exh = sE[cur][0][0][tx + 1];
            const float vx = cu.hh0 - chy * (sE[cur][2][1][tx + 1] - ezh) + chy * (sE[nxt][1][0][tx + 1] - eyh);
            const float vz = cu.hh1 - chx * (sE[cur][1][0][tx + 2] - eyh) + chy * (sE[cur][0][1][tx + 1] - exh);
            sH[0][0][tx + 1] = (rx_ij && kh) ? vx : cu.hh0;
            sH[2][0][tx + 1] = (rz_ij && kz) ? vz : cu.hh1;
        }
        if (w_col) {  // (i0-1, j0-1+ccr): Hy, Hz
            const float exh = sE[cur][0][ccr][0], eyh = sE[cur][1][ccr][0], ezh = sE[cur][2][ccr][0];
            const float vy = cu.hh0 - chz * (sE[nxt][0][ccr][0] - exh) + chx * (sE[cur][2][ccr][1] - ezh);
            const float vz = cu.hh1 - chx * (sE[cur][1][ccr][1] - eyh) + chy * (sE[cur][0][ccr + 1][0] - exh);
            *col_dst_y = (cy_ij && kh) ? vy : cu.hh0;
            *col_dst_z = (cz_ij && kz) ? vz : cu.hh1;
        }
        const uint32_t ob = (uint32_t)k * szb + c_own;
        if (own) {
            X::st_nt(hxn, ob, hx); X::st_nt(hyn, ob, hy); X::st_nt(hzn, ob, hz);
        }
        __syncthreads();  // H^{n+1/2} of plane k and its halo staged

        // A source cell's Dz is written even outside the Dz update range (the
        // upper x/y faces and the top plane), as updateSources does
        // (fdtd_calculations.cu:187-211).
        const bool rdx = dx_ij && kz, rdy = dy_ij && kz, rdz = (dz_ij && kh) || (cu.id & kSourceBit);
        const bool reyz = e_i && (k >= 1);
        const float dxu = cu.dxn + cdy * (hz - sH[2][ty][tx + 1]) - cdz * (hy - hym);
        const float dyu = cu.dyn + cdz * (hx - hxm) - cdx * (hz - sH[2][ty + 1][tx]);
        float dzu = cu.dzn + cdx * (hy - sH[1][ty + 1][tx]) - cdy * (hx - sH[0][ty][tx + 1]);
        dzu = (cu.id & kSourceBit) ? dzu - jzv : dzu;
        float dyp = dyu, dzp = dzu;
        // E cells whose D is not updated read the target slot's value.
        if (own && reyz && !rdy) dyp = X::ld(dyT, ob);
        if (own && reyz && !rdz) dzp = X::ld(dzT, ob);
        const int m = cu.id & kMaterialMask;
        float exn, eyn, ezn;
        if constexpr (exact) {
            const float ca = lA[m], cc = lC[m], cd = lD[m];
            const double g1 = lG1[m], g2 = lG2[m], g3 = lG3[m];
            if (ZS && cu.zs) {
                // The skipped terms are a*E^{n-1} and g3*D^{n-1} with a = g3 = +0, so the
                // result is the reference's unless d*E^n is -0 (a zero sum then takes the
                // sign of the dropped term, which may be -0) or E^n or D^n is not finite
                // (E^{n-1} or D^{n-1} can only be infinite or NaN where they are, and
                // 0*inf is NaN).  If any skipping lane of the wave holds such a cell, every
                // skipping lane loads the real operands after all.
                auto neg0 = [](float v) { return __float_as_uint(v) == 0x80000000u; };
                auto fin = [](float v) { return __builtin_isfinite(v); };
                const bool risky = own && (neg0(cd * exk) || neg0(cd * eyk) || neg0(cd * ezk) || !fin(exk) ||
                                           !fin(eyk) || !fin(ezk) || !fin(cu.dxn) || !fin(cu.dyn) || !fin(cu.dzn));
                if (__any(risky)) {
                    cu.exm = X::ld_nt(ex1, ob); cu.eym = X::ld_nt(ey1, ob); cu.ezm = X::ld_nt(ez1, ob);
                    cu.dxm = X::ld_nt(dx1, ob); cu.dym = X::ld_nt(dy1, ob); cu.dzm = X::ld_nt(dz1, ob);
                }
            }
            exn = e_update(cc, cd, ca, g1, g2, g3, exk, cu.exm, dxu, cu.dxn, cu.dxm);
            eyn = e_update(cc, cd, ca, g1, g2, g3, eyk, cu.eym, dyp, cu.dyn, cu.dym);
            ezn = e_update(cc, cd, ca, g1, g2, g3, ezk, cu.ezm, dzp, cu.dzn, cu.dzm);
        } else {
            const float k1 = lK1[m], k2 = lK2[m], k3 = lK3[m], k5 = lK5[m];
            exn = e_update_scaled(k1, k2, k3, k5, exk, cu.exm, dxu, cu.dxn, cu.dxm);
            eyn = e_update_scaled(k1, k2, k3, k5, eyk, cu.eym, dyp, cu.dyn, cu.dym);
            ezn = e_update_scaled(k1, k2, k3, k5, ezk, cu.ezm, dzp, cu.dzn, cu.dzm);
        }
        // Mur x faces (fdtd_calculations.cu:303-329, 369-395) in the tile:
        // E^{n+1} of the neighbour comes from the adjacent lane as it will be
        // stored, E^n of the face and the neighbour from LDS.  Ez at k = 0 is
        // outside the E update, so its neighbour is read from the target
        // slot (not written in this step).
        bool sty = reyz, stz = reyz;
        if (mur_lo) {
            const float ey1 = __shfl_down(X::rnd(eyn), 1, 64), ez1 = __shfl_down(X::rnd(ezn), 1, 64);
            if (tx == 0 && j < ny) {
                const float4 c = sMurC[0][k - k0][ty];
                if (j < ny - 1 && k >= 1 && k < nz - 1) {
                    eyn = c.x * ((c.y * ey1 + c.z * sE[cur][1][ty + 1][2]) - c.w * eyk);
                    sty = true;
                }
                if (j >= 1 && j < ny - 1 && k < nz - 1) {
                    const float tn = k >= 1 ? ez1 : sMurT[0][ty];
                    ezn = c.x * ((c.y * tn + c.z * sE[cur][2][ty + 1][2]) - c.w * ezk);
                    stz = true;
                }
            }
        }
        if (mur_hi) {
            const float ezm = __shfl_up(X::rnd(ezn), 1, 64);
            if (i == nx - 1 && j < ny) {
                const float4 c = sMurC[1][k - k0][ty];
                if (j < ny - 1 && k >= 1 && k < nz - 1) {
                    // the x-high Ey face reads the source neighbour twice (fdtd_calculations.cu:323-325)
                    const float sn = sE[cur][1][ty + 1][tx];
                    eyn = c.x * ((c.y * sn + c.z * sn) - c.w * eyk);
                    sty = true;
                }
                if (j >= 1 && j < ny - 1 && k < nz - 1) {
                    const float tn = k >= 1 ? ezm : sMurT[1][ty];
                    ezn = c.x * ((c.y * tn + c.z * sE[cur][2][ty + 1][tx]) - c.w * ezk);
                    stz = true;
                }
            }
        }
        if (own) {
            if (rdx) { X::st_nt(dxT, ob, dxu); X::st_nt(exT, ob, exn); }
            if (rdy) X::st_nt(dyT, ob, dyu);
            if (rdz) X::st_nt(dzT, ob, dzu);
            if (sty) X::st_nt(eyT, ob, eyn);
            if (stz) X::st_nt(ezT, ob, ezn);
        }
        hxm = hx;
        hym = hy;
        cur = nxt;
        if (more) cu = nx_;
        if constexpr (ZS) zsn = zskip && k + 2 < k1 && zero_row(idq);   // for plane k+2
    }
}

// 16-bit pairs: one 32-bit word holds the values of cells (2m, 2m+1) of a row.
template <typename T> struct Pair;
template <> struct Pair<__half> {
    static __device__ __forceinline__ float lo(uint32_t w) { return Elem<__half>::cvt((unsigned short)(w & 0xffffu)); }
    static __device__ __forceinline__ float hi(uint32_t w) { return Elem<__half>::cvt((unsigned short)(w >> 16)); }
    static __device__ __forceinline__ uint32_t pack(float a, float b)
    {
        return (uint32_t)Elem<__half>::enc(a) | ((uint32_t)Elem<__half>::enc(b) << 16);
    }
};
template <> struct Pair<__hip_bfloat16> {
    static __device__ __forceinline__ float lo(uint32_t w) { return __uint_as_float(w << 16); }
    static __device__ __forceinline__ float hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
    static __device__ __forceinline__ uint32_t pack(float a, float b)
    {
        return (uint32_t)Elem<__hip_bfloat16>::enc(a) | ((uint32_t)Elem<__hip_bfloat16>::enc(b) << 16);
    }
};

__device__ __forceinline__ uint32_t ldw(const void *p, uint32_t b) { return *(const uint32_t *)((const char *)p + b); }
__device__ __forceinline__ uint32_t ldw_nt(const void *p, uint32_t b)
{
    return __builtin_nontemporal_load((const uint32_t *)((const char *)p + b));
}
__device__ __forceinline__ void stw_nt(void *p, uint32_t b, uint32_t w)
{
    __builtin_nontemporal_store(w, (uint32_t *)((char *)p + b));
}

// The fused step of one 128 x TY x kchunk tile for 16-bit storage, two
// cells per lane: lane l of a wave holds cells (i0 + 2l, i0 + 2l + 1) of its
// row, so every stream moves one 32-bit word per lane (256 bytes per wave
// access) instead of 16 bits.  Each cell's arithmetic is that of
// fused_tile<TY, T, false>, operation for operation, so the two kernels give
// identical results.  Pairs are split into 16-bit stores only where the two
// cells' update ranges differ (the x boundaries and the padding).
// LDS row of a paired tile, split into even and odd cells: column c of the
// unsplit row (cell x = c - 1 of the tile, x in [-1, 2 * TX]) is half x & 1,
// index (x >> 1) + 1.
template <int N>
__device__ __forceinline__ float &lx(float (&row)[2][N], int c)
{
    const int x = c - 1;
    return row[x & 1][(x >> 1) + 1];
}

template <int TY, typename T>
__device__ __forceinline__ void fused_tile_pair(const StepArgs &A, int bx, int by, int bz)
{
    static_assert(sizeof(T) == 2, "pairs of 16-bit cells");
    static_assert(TY >= 2 && TY <= 8, "tile rows");
    constexpr int TC = 2 * TX;  // cells per tile row
    constexpr uint32_t ES = 2;
    using X = Elem<T>;
    using P = Pair<T>;
    // Rows hold the even cells (x = 2m) and the odd cells (x = 2m + 1) in two
    // halves (lx), so the two cells of lane tx sit at index tx + 1 of each
    // half and every lane-indexed access is unit-stride: no bank conflicts.
    __shared__ float sE[2][3][TY + 2][2][TX + 2];
    __shared__ float sH[3][TY + 1][2][TX + 1];
    __shared__ float sJunk[2][TX];
    __shared__ float4 sMurC[2][kMurK][TY];
    __shared__ float sMurT[2][TY];

    // every wave covers one tile row (64 lanes along x), so the row index is
    // wave-uniform: as a scalar it keeps the row tests and the wave roles
    // (halo rows, column gather) out of the per-lane masks
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int i0 = bx * TC, j0 = by * TY;
    const int k0 = bz * A.kchunk;
    const int k1 = min(k0 + A.kchunk, nz);
    const int ia = i0 + 2 * tx;  // cells ia (a) and ia + 1 (b)
    const int j = j0 + ty;
    const int ca = 2 * tx + 1;   // LDS column of cell a; b is ca + 1
    const bool own_a = ia < nx && j < ny, own_b = ia + 1 < nx && j < ny;

    const int nmat = A.nmat;
    float *lK1 = (float *)lut_lds(), *lK2 = lK1 + nmat, *lK3 = lK1 + 2 * nmat, *lK5 = lK1 + 3 * nmat;

    const uint32_t gb = A.guard * ES;
    auto sh = [&](const void *p) { return (const T *)((const char *)p - gb); };
    auto shw = [&](void *p) { return (T *)((char *)p - gb); };
    const T *hxo = sh(A.hxo), *hyo = sh(A.hyo), *hzo = sh(A.hzo);
    const T *ex0 = sh(A.ex0), *ey0 = sh(A.ey0), *ez0 = sh(A.ez0);
    const T *ex1 = sh(A.ex1), *ey1 = sh(A.ey1), *ez1 = sh(A.ez1);
    const T *dx0 = sh(A.dx0), *dy0 = sh(A.dy0), *dz0 = sh(A.dz0);
    const T *dx1 = sh(A.dx1), *dy1 = sh(A.dy1), *dz1 = sh(A.dz1);
    T *hxn = shw(A.hxn), *hyn = shw(A.hyn), *hzn = shw(A.hzn);
    T *exT = shw(A.exT), *eyT = shw(A.eyT), *ezT = shw(A.ezT);
    T *dxT = shw(A.dxT), *dyT = shw(A.dyT), *dzT = shw(A.dzT);
    const uint8_t *ids = A.ids - A.guard;

    const uint32_t syb = (uint32_t)A.pitch * ES, szb = (uint32_t)A.plane * ES;
    // Pairs past the row pitch read the last pair of the row (unused values).
    const int il = min(ia, A.pitch - 2);
    const uint32_t c_own = gb + (uint32_t)j * syb + (uint32_t)il * ES;
    const uint32_t c_rlo = gb + (uint32_t)(j0 - 1) * syb + (uint32_t)il * ES;
    const uint32_t c_rhi = gb + (uint32_t)(j0 + TY) * syb + (uint32_t)il * ES;
    const int side = tx >> 5;
    const int cr = min(tx & 31, TY + 1);
    const int cl = side ? TC + 1 : 0;
    const int ci = side ? i0 + TC : i0 - 1;
    const uint32_t c_col = gb + (uint32_t)(j0 - 1 + cr) * syb + (uint32_t)ci * ES;
    const bool w_rlo = (ty == 0), w_rhi = (ty == TY - 1), w_col = (ty == 1);

    const float chx = A.chx, chy = A.chy, chz = A.chz;
    const float cdx = A.cdx, cdy = A.cdy, cdz = A.cdz;

    // E^n staging: the lane's pair (word), -y row pair (wave 0), +y row pair
    // of Ex and Ez (wave TY-1), gathered column cells (wave 1).
    uint32_t pe[3], pr[3], pq[2];
    float pc[3];
    auto fetch_e = [&](int k) {
        const uint32_t kb = (uint32_t)k * szb;
        pe[0] = ldw(ex0, kb + c_own); pe[1] = ldw(ey0, kb + c_own); pe[2] = ldw(ez0, kb + c_own);
        if (w_rlo) { pr[0] = ldw(ex0, kb + c_rlo); pr[1] = ldw(ey0, kb + c_rlo); pr[2] = ldw(ez0, kb + c_rlo); }
        if (w_rhi) { pq[0] = ldw(ex0, kb + c_rhi); pq[1] = ldw(ez0, kb + c_rhi); }
        if (w_col) { pc[0] = X::ld(ex0, kb + c_col); pc[1] = X::ld(ey0, kb + c_col); pc[2] = X::ld(ez0, kb + c_col); }
    };
    auto stage_e = [&](int buf) {
        for (int c = 0; c < 3; c++) {
            lx(sE[buf][c][ty + 1], ca) = P::lo(pe[c]);
            lx(sE[buf][c][ty + 1], ca + 1) = P::hi(pe[c]);
        }
        if (w_rlo)
            for (int c = 0; c < 3; c++) {
                lx(sE[buf][c][0], ca) = P::lo(pr[c]);
                lx(sE[buf][c][0], ca + 1) = P::hi(pr[c]);
            }
        if (w_rhi) {
            lx(sE[buf][0][TY + 1], ca) = P::lo(pq[0]); lx(sE[buf][0][TY + 1], ca + 1) = P::hi(pq[0]);
            lx(sE[buf][2][TY + 1], ca) = P::lo(pq[1]); lx(sE[buf][2][TY + 1], ca + 1) = P::hi(pq[1]);
        }
        if (w_col) { lx(sE[buf][0][cr], cl) = pc[0]; lx(sE[buf][1][cr], cl) = pc[1]; lx(sE[buf][2][cr], cl) = pc[2]; }
    };

    // Own-pair operands of one plane (words), two halo words, the id pair.
    struct Own {
        uint32_t hxo, hyo, hzo, dxn, dyn, dzn, dxm, dym, dzm, exm, eym, ezm, hh0, hh1;
        uint32_t id;
    };
    auto fetch_own = [&](int k, Own &o) {
        const uint32_t ob = (uint32_t)k * szb + c_own;
        o.hxo = ldw(hxo, ob); o.hyo = ldw(hyo, ob); o.hzo = ldw(hzo, ob);
        o.dxn = ldw_nt(dx0, ob); o.dyn = ldw_nt(dy0, ob); o.dzn = ldw_nt(dz0, ob);
        o.dxm = ldw_nt(dx1, ob); o.dym = ldw_nt(dy1, ob); o.dzm = ldw_nt(dz1, ob);
        o.exm = ldw_nt(ex1, ob); o.eym = ldw_nt(ey1, ob); o.ezm = ldw_nt(ez1, ob);
        o.id = __builtin_nontemporal_load((const unsigned short *)(ids + ob / ES));
        // halo H^{n-1/2}: row pair (ia, j0-1): Hx, Hz words; column cell (i0-1 | i0+TC, row cr): Hy, Hz
        o.hh0 = 0u; o.hh1 = 0u;
        if (w_rlo) { const uint32_t oh = (uint32_t)k * szb + c_rlo; o.hh0 = ldw(hxo, oh); o.hh1 = ldw(hzo, oh); }
        if (w_col) {
            const uint32_t oh = (uint32_t)k * szb + c_col;
            o.hh0 = __float_as_uint(X::ld(hyo, oh));
            o.hh1 = __float_as_uint(X::ld(hzo, oh));
        }
    };
    auto half_of = [](uint32_t w, int h) { return h ? P::hi(w) : P::lo(w); };

    // Mur x faces, staged as in fused_tile.
    const bool mur_lo = A.mur_x && i0 == 0;
    const bool mur_hi = A.mur_x == 1 && i0 <= nx - 1 && nx - 1 < i0 + TC;
    if (mur_lo || mur_hi) {
        const int tid = ty * TX + tx, nk = k1 - k0;
        for (int q = tid; q < 2 * nk * TY; q += TX * TY) {
            const int sd = q / (nk * TY), r = q - sd * nk * TY, kk = r / TY, row = r - kk * TY;
            const long long at = (long long)(k0 + kk) * ny + min(j0 + row, ny - 1);
            sMurC[sd][kk][row] = sd ? A.murx1[at] : A.murx0[at];
        }
        if (k0 == 0 && tid < 2 * TY) {
            const int sd = tid / TY, row = tid - sd * TY;
            const int ii = sd ? nx - 2 : 1, jj = min(j0 + row, ny - 1);
            sMurT[sd][row] = X::ld(ezT, gb + (uint32_t)jj * syb + (uint32_t)ii * ES);
        }
    }

    float hxm[2] = {0.f, 0.f}, hym[2] = {0.f, 0.f};  // own-column Hx, Hy of plane k-1
    int cur = 0;
    if (k0 > 0) {
        fetch_e(k0 - 1);
        stage_e(0);
        fetch_e(k0);
        stage_e(1);
        __syncthreads();
        const int k = k0 - 1;
        const uint32_t ob = (uint32_t)k * szb + c_own;
        const uint32_t hx0 = ldw(hxo, ob), hy0 = ldw(hyo, ob);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = ca + h, i = ia + h;
            const float hxv = half_of(hx0, h), hyv = half_of(hy0, h);
            const float exk = lx(sE[0][0][ty + 1], c), eyk = lx(sE[0][1][ty + 1], c), ezk = lx(sE[0][2][ty + 1], c);
            const float exk1 = lx(sE[1][0][ty + 1], c), eyk1 = lx(sE[1][1][ty + 1], c);
            const float ux = hxv - chy * (lx(sE[0][2][ty + 2], c) - ezk) + chy * (eyk1 - eyk);
            const float uy = hyv - chz * (exk1 - exk) + chx * (lx(sE[0][2][ty + 1], c + 1) - ezk);
            hxm[h] = (i >= 1 && i < nx - 1 && j < ny - 1 && k < nz - 1) ? ux : hxv;
            hym[h] = (i < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1) ? uy : hyv;
        }
        cur = 1;
        __syncthreads();
    } else {
        fetch_e(k0);
        stage_e(0);
    }
    Own cu, nx_;
    fetch_own(k0, cu);
    fetch_e(k0 + 1);

    const int jh = j0 - 1;
    const int ccr = min(max(cr, 1), TY);
    const int jc = j0 - 1 + ccr;
    const bool col_ok = (side == 0) && (cr >= 1) && (cr <= TY) && (i0 >= 1);
    const bool cy_ij = col_ok && (i0 - 1 < nx - 1) && (jc >= 1 && jc < ny - 1);
    const bool cz_ij = col_ok && (i0 - 1 < nx - 1) && (jc < ny - 1);
    float *col_dst_y = col_ok ? &lx(sH[1][ccr], 0) : &sJunk[0][tx];
    float *col_dst_z = col_ok ? &lx(sH[2][ccr], 0) : &sJunk[1][tx];

    const float jzv = A.jz;
    for (int k = k0; k < k1; ++k) {
        const int nxt = cur ^ 1;
        stage_e(nxt);
        const bool more = (k + 1 < k1);
        if (more) {
            fetch_own(k + 1, nx_);
            fetch_e(k + 2);
        }
        __syncthreads();

        const bool kh = (k < nz - 1), kz = (k >= 1 && k < nz - 1);
        float hx[2], hy[2], hz[2], exk[2], eyk[2], ezk[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = ca + h, i = ia + h;
            exk[h] = lx(sE[cur][0][ty + 1], c); eyk[h] = lx(sE[cur][1][ty + 1], c); ezk[h] = lx(sE[cur][2][ty + 1], c);
            const float exk1 = lx(sE[nxt][0][ty + 1], c), eyk1 = lx(sE[nxt][1][ty + 1], c);
            const float hxo_ = half_of(cu.hxo, h), hyo_ = half_of(cu.hyo, h), hzo_ = half_of(cu.hzo, h);
            const float ux = hxo_ - chy * (lx(sE[cur][2][ty + 2], c) - ezk[h]) + chy * (eyk1 - eyk[h]);
            const float uy = hyo_ - chz * (exk1 - exk[h]) + chx * (lx(sE[cur][2][ty + 1], c + 1) - ezk[h]);
            const float uz = hzo_ - chx * (lx(sE[cur][1][ty + 1], c + 1) - eyk[h]) + chy * (lx(sE[cur][0][ty + 2], c) - exk[h]);
            hx[h] = (i >= 1 && i < nx - 1 && j < ny - 1 && kh) ? ux : hxo_;
            hy[h] = (i < nx - 1 && j >= 1 && j < ny - 1 && kh) ? uy : hyo_;
            hz[h] = (i < nx - 1 && j < ny - 1 && kz) ? uz : hzo_;
            lx(sH[0][ty + 1], c) = hx[h];
            lx(sH[1][ty + 1], c) = hy[h];
            lx(sH[2][ty + 1], c) = hz[h];
        }
        if (w_rlo) {  // row pair (ia + h, j0-1): Hx, Hz
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int c = ca + h, i = ia + h;
                const float h0 = half_of(cu.hh0, h), h1 = half_of(cu.hh1, h);
                const float ezh = lx(sE[cur][2][0], c), eyh = lx(sE[cur][1][0], c), exh = lx(sE[cur][0][0], c);
                const float vx = h0 - chy * (lx(sE[cur][2][1], c) - ezh) + chy * (lx(sE[nxt][1][0], c) - eyh);
                const float vz = h1 - chx * (lx(sE[cur][1][0], c + 1) - eyh) + chy * (lx(sE[cur][0][1], c) - exh);
                const bool rx = (jh >= 0) && (i >= 1 && i < nx - 1 && jh < ny - 1);
                const bool rz = (jh >= 0) && (i < nx - 1 && jh < ny - 1);
                lx(sH[0][0], c) = (rx && kh) ? vx : h0;
                lx(sH[2][0], c) = (rz && kz) ? vz : h1;
            }
        }
        if (w_col) {  // (i0-1, j0-1+ccr): Hy, Hz
            const float h0 = __uint_as_float(cu.hh0), h1 = __uint_as_float(cu.hh1);
            const float exh = lx(sE[cur][0][ccr], 0), eyh = lx(sE[cur][1][ccr], 0), ezh = lx(sE[cur][2][ccr], 0);
            const float vy = h0 - chz * (lx(sE[nxt][0][ccr], 0) - exh) + chx * (lx(sE[cur][2][ccr], 1) - ezh);
            const float vz = h1 - chx * (lx(sE[cur][1][ccr], 1) - eyh) + chy * (lx(sE[cur][0][ccr + 1], 0) - exh);
            *col_dst_y = (cy_ij && kh) ? vy : h0;
            *col_dst_z = (cz_ij && kz) ? vz : h1;
        }
        const uint32_t ob = (uint32_t)k * szb + c_own;
        // one store of a pair, or of its first cell alone at the row's end
        auto st_pair = [&](T *p, bool sa, bool sb, float va, float vb) {
            if (sa && sb) stw_nt(p, ob, P::pack(va, vb));
            else if (sa) X::st_nt(p, ob, va);
            else if (sb) X::st_nt(p, ob + ES, vb);
        };
        st_pair(hxn, own_a, own_b, hx[0], hx[1]);
        st_pair(hyn, own_a, own_b, hy[0], hy[1]);
        st_pair(hzn, own_a, own_b, hz[0], hz[1]);
        __syncthreads();

        float dxu[2], dyu[2], dzu[2], exn[2], eyn[2], ezn[2];
        bool rdx[2], rdy[2], rdz[2], reyz[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = ca + h, i = ia + h;
            const bool own = h ? own_b : own_a;
            rdx[h] = (i < nx - 1 && j >= 1 && j < ny - 1) && kz;
            rdy[h] = (i >= 1 && i < nx - 1 && j < ny - 1) && kz;
            const uint32_t id = (cu.id >> (8 * h)) & 0xffu;
            rdz[h] = ((i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1) && kh) || (id & kSourceBit);
            reyz[h] = (i >= 1) && (k >= 1);
            const float dxn_ = half_of(cu.dxn, h), dyn_ = half_of(cu.dyn, h), dzn_ = half_of(cu.dzn, h);
            dxu[h] = dxn_ + cdy * (hz[h] - lx(sH[2][ty], c)) - cdz * (hy[h] - hym[h]);
            dyu[h] = dyn_ + cdz * (hx[h] - hxm[h]) - cdx * (hz[h] - lx(sH[2][ty + 1], c - 1));
            float dz_ = dzn_ + cdx * (hy[h] - lx(sH[1][ty + 1], c - 1)) - cdy * (hx[h] - lx(sH[0][ty], c));
            dz_ = (id & kSourceBit) ? dz_ - jzv : dz_;
            dzu[h] = dz_;
            float dyp = dyu[h], dzp = dz_;
            if (own && reyz[h] && !rdy[h]) dyp = X::ld(dyT, ob + h * ES);
            if (own && reyz[h] && !rdz[h]) dzp = X::ld(dzT, ob + h * ES);
            const int m = id & kMaterialMask;
            const float k1c = lK1[m], k2c = lK2[m], k3c = lK3[m], k5c = lK5[m];
            exn[h] = e_update_scaled(k1c, k2c, k3c, k5c, exk[h], half_of(cu.exm, h), dxu[h], dxn_, half_of(cu.dxm, h));
            eyn[h] = e_update_scaled(k1c, k2c, k3c, k5c, eyk[h], half_of(cu.eym, h), dyp, dyn_, half_of(cu.dym, h));
            ezn[h] = e_update_scaled(k1c, k2c, k3c, k5c, ezk[h], half_of(cu.ezm, h), dzp, dzn_, half_of(cu.dzm, h));
        }
        bool sty[2] = {reyz[0], reyz[1]}, stz[2] = {reyz[0], reyz[1]};
        if (mur_lo && tx == 0 && j < ny) {  // face i = 0 is cell a of lane 0, its neighbour cell b
            const float4 cm = sMurC[0][k - k0][ty];
            if (j < ny - 1 && k >= 1 && k < nz - 1) {
                eyn[0] = cm.x * ((cm.y * X::rnd(eyn[1]) + cm.z * lx(sE[cur][1][ty + 1], 2)) - cm.w * eyk[0]);
                sty[0] = true;
            }
            if (j >= 1 && j < ny - 1 && k < nz - 1) {
                const float tn = k >= 1 ? X::rnd(ezn[1]) : sMurT[0][ty];
                ezn[0] = cm.x * ((cm.y * tn + cm.z * lx(sE[cur][2][ty + 1], 2)) - cm.w * ezk[0]);
                stz[0] = true;
            }
        }
        if (mur_hi) {
            // face i = nx-1 is cell b of the lane holding nx-2 (odd offset) or
            // cell a with the neighbour in the previous lane's cell b
            const bool face_b = ((nx - 1 - i0) & 1) != 0;
            const float ezm = __shfl_up(X::rnd(ezn[1]), 1, 64);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (ia + h != nx - 1 || j >= ny || (h == 1) != face_b) continue;
                const float4 cm = sMurC[1][k - k0][ty];
                const int c = ca + h;
                if (j < ny - 1 && k >= 1 && k < nz - 1) {
                    const float sn = lx(sE[cur][1][ty + 1], c - 1);
                    eyn[h] = cm.x * ((cm.y * sn + cm.z * sn) - cm.w * eyk[h]);
                    sty[h] = true;
                }
                if (j >= 1 && j < ny - 1 && k < nz - 1) {
                    const float tn = k >= 1 ? (h ? X::rnd(ezn[0]) : ezm) : sMurT[1][ty];
                    ezn[h] = cm.x * ((cm.y * tn + cm.z * lx(sE[cur][2][ty + 1], c - 1)) - cm.w * ezk[h]);
                    stz[h] = true;
                }
            }
        }
        st_pair(dxT, own_a && rdx[0], own_b && rdx[1], dxu[0], dxu[1]);
        st_pair(exT, own_a && rdx[0], own_b && rdx[1], exn[0], exn[1]);
        st_pair(dyT, own_a && rdy[0], own_b && rdy[1], dyu[0], dyu[1]);
        st_pair(dzT, own_a && rdz[0], own_b && rdz[1], dzu[0], dzu[1]);
        st_pair(eyT, own_a && sty[0], own_b && sty[1], eyn[0], eyn[1]);
        st_pair(ezT, own_a && stz[0], own_b && stz[1], ezn[0], ezn[1]);
        hxm[0] = hx[0]; hxm[1] = hx[1];
        hym[0] = hy[0]; hym[1] = hy[1];
        cur = nxt;
        if (more) cu = nx_;
    }
}

// One grid, one step: block b runs tile xcd_logical(b), at five waves per
// SIMD (96 VGPRs).
template <int TY, typename T, bool EXACT, int TXW, bool ZS = false>
__global__ __launch_bounds__(TXW *TY) __attribute__((amdgpu_waves_per_eu(5))) void fused_step(const StepArgs A)
{
    const int tid = threadIdx.y * TXW + threadIdx.x;
    const int nblk = A.tiles_x * A.tiles_y * A.tiles_z;
    const int b = xcd_logical((int)blockIdx.x, nblk, A.xcd_remap, xcd_unit(A.xcd_remap, A.tiles_x, A.tiles_y));
    const int bx = b % A.tiles_x;
    const int by = (b / A.tiles_x) % A.tiles_y;
    const int bz = b / (A.tiles_x * A.tiles_y);
    stage_lut<EXACT>(A.lut, A.nmat, tid, TXW * TY);
    fused_tile<TY, T, EXACT, TXW, ZS>(A, bx, by, bz);
}

// Two time steps per launch (fp32 storage, the reference's arithmetic).
//
// Step n -> n+1 runs on a 64 x R box of (x, y) columns at origin
// (62 bx - 1, (R-2) by - 1): one wave per box row, one cell per lane, with
// the E^n staging, halo H and z-march of fused_tile.  Step n+1 -> n+2 needs
// E^{n+1} one cell around each of its cells, so a tile stores only the
// inner 62 x (R-2) cells of its box, and the boxes of neighbouring tiles
// overlap by two columns and two rows.  Both steps march z together, the
// second one plane behind the first: iteration kA computes step n -> n+1
// on plane kA and step n+1 -> n+2 on plane kA-1.  Between the two, the first
// step's six Mur faces (fdtd_calculations.cu:216-428) are applied inside the
// box, so the second step reads E^{n+1} as the reference's step leaves it;
// the second step's faces are left to the Mur kernel.  H^{n+1/2} stays on
// the chip.  Step n -> n+1 reads E^{n-1} and D^{n-1} on every box cell,
// including the ring that neighbouring tiles store, so step n+1 -> n+2
// writes level n+2 into a spare slot rather than over level n-1.
// Per cell and launch: 61 bytes read (H^{n-1/2}, E^n, E^{n-1}, D^n, D^{n-1},
// the id) and 60 written (H^{n+3/2}, E and D of levels n+1 and n+2).
template <int R>
__device__ __forceinline__ void fused2_tile(const Step2Args &S, int bx, int by, int bz)
{
    // The +y halo row (j0 + R <= ny + R - 2) and plane nz stay inside the
    // arrays' guard of one plane and 10 rows.
    static_assert(R >= 4 && R <= 12, "tile rows");
    constexpr int W = 64;
    constexpr uint32_t ES = sizeof(float);
    using X = Elem<float>;
    const StepArgs &A = S.A;
    __shared__ float sE[2][3][R + 2][W + 2];   // E^n of planes kA and kA+1, one-cell halo
    __shared__ float sH[3][R + 1][W + 1];      // H^{n+1/2} of plane kA with its -x/-y halo
    __shared__ float sF[2][3][R + 1][W + 1];   // E^{n+1} of planes kA-1 and kA (box; padding row and column)
    __shared__ float sG[3][R + 1][W + 1];      // H^{n+3/2} of plane kA-1 (box at [1..R][1..W])
    __shared__ float sJunk[2][W];

    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int i0 = bx * (W - 2) - 1, j0 = by * (R - 2) - 1;
    const int k0 = bz * A.kchunk;
    const int k1 = min(k0 + A.kchunk, nz);
    const int i = i0 + tx, j = j0 + ty;
    const bool in = i >= 0 && i < nx && j >= 0 && j < ny;
    const bool own = in && tx >= 1 && tx <= W - 2 && ty >= 1 && ty <= R - 2;   // the cells this tile stores

    const int nmat = A.nmat;
    const double *sLut = lut_lds();
    const double *lG1 = sLut, *lG2 = sLut + nmat, *lG3 = sLut + 2 * nmat;
    const float *lA = (const float *)(sLut + 3 * nmat), *lC = lA + nmat, *lD = lA + 2 * nmat;

    const uint32_t gb = A.guard * ES;
    auto sh = [&](const void *p) { return (const float *)((const char *)p - gb); };
    auto shw = [&](void *p) { return (float *)((char *)p - gb); };
    const float *hxo = sh(A.hxo), *hyo = sh(A.hyo), *hzo = sh(A.hzo);
    const float *ex0 = sh(A.ex0), *ey0 = sh(A.ey0), *ez0 = sh(A.ez0);
    const float *ex1 = sh(A.ex1), *ey1 = sh(A.ey1), *ez1 = sh(A.ez1);
    const float *dx0 = sh(A.dx0), *dy0 = sh(A.dy0), *dz0 = sh(A.dz0);
    const float *dx1 = sh(A.dx1), *dy1 = sh(A.dy1), *dz1 = sh(A.dz1);
    float *hxn = shw(A.hxn), *hyn = shw(A.hyn), *hzn = shw(A.hzn);
    float *exT = shw(A.exT), *eyT = shw(A.eyT), *ezT = shw(A.ezT);
    float *dxT = shw(A.dxT), *dyT = shw(A.dyT), *dzT = shw(A.dzT);
    float *exS = shw(S.exS), *eyS = shw(S.eyS), *ezS = shw(S.ezS);
    float *dxS = shw(S.dxS), *dyS = shw(S.dyS), *dzS = shw(S.dzS);
    const uint8_t *ids = A.ids - A.guard;

    const uint32_t syb = (uint32_t)A.pitch * ES, szb = (uint32_t)A.plane * ES;
    // 32-bit offsets wrap for the box's -1 row and column (and the -2 halo).
    const int il = min(i, A.pitch - 1);
    const uint32_t c_own = gb + (uint32_t)j * syb + (uint32_t)il * ES;
    const uint32_t c_rlo = gb + (uint32_t)(j0 - 1) * syb + (uint32_t)il * ES;
    const uint32_t c_rhi = gb + (uint32_t)(j0 + R) * syb + (uint32_t)il * ES;
    const int side = tx >> 5;
    const int cr = min(tx & 31, R + 1);
    const int cl = side ? W + 1 : 0;
    const int ci = side ? i0 + W : i0 - 1;
    const uint32_t c_col = gb + (uint32_t)(j0 - 1 + cr) * syb + (uint32_t)ci * ES;
    const bool w_rlo = (ty == 0), w_rhi = (ty == R - 1), w_col = (ty == 1);

    const float chx = A.chx, chy = A.chy, chz = A.chz;
    const float cdx = A.cdx, cdy = A.cdy, cdz = A.cdz;

    float pe[3], pr[3], pq[2], pc[3];
    auto fetch_e = [&](int k) {
        const uint32_t kb = (uint32_t)k * szb;
        pe[0] = X::ld(ex0, kb + c_own); pe[1] = X::ld(ey0, kb + c_own); pe[2] = X::ld(ez0, kb + c_own);
        if (w_rlo) { pr[0] = X::ld(ex0, kb + c_rlo); pr[1] = X::ld(ey0, kb + c_rlo); pr[2] = X::ld(ez0, kb + c_rlo); }
        if (w_rhi) { pq[0] = X::ld(ex0, kb + c_rhi); pq[1] = X::ld(ez0, kb + c_rhi); }
        if (w_col) { pc[0] = X::ld(ex0, kb + c_col); pc[1] = X::ld(ey0, kb + c_col); pc[2] = X::ld(ez0, kb + c_col); }
    };
    auto stage_e = [&](int buf) {
        sE[buf][0][ty + 1][tx + 1] = pe[0]; sE[buf][1][ty + 1][tx + 1] = pe[1]; sE[buf][2][ty + 1][tx + 1] = pe[2];
        if (w_rlo) { sE[buf][0][0][tx + 1] = pr[0]; sE[buf][1][0][tx + 1] = pr[1]; sE[buf][2][0][tx + 1] = pr[2]; }
        if (w_rhi) { sE[buf][0][R + 1][tx + 1] = pq[0]; sE[buf][2][R + 1][tx + 1] = pq[1]; }
        if (w_col) { sE[buf][0][cr][cl] = pc[0]; sE[buf][1][cr][cl] = pc[1]; sE[buf][2][cr][cl] = pc[2]; }
    };
    struct Own {
        float hxo, hyo, hzo, dxn, dyn, dzn, dxm, dym, dzm, exm, eym, ezm, hh0, hh1;
        uint32_t id;
    };
    auto fetch_own = [&](int k, Own &o) {
        const uint32_t ob = (uint32_t)k * szb + c_own;
        o.hxo = X::ld(hxo, ob); o.hyo = X::ld(hyo, ob); o.hzo = X::ld(hzo, ob);
        o.dxn = X::ld_nt(dx0, ob); o.dyn = X::ld_nt(dy0, ob); o.dzn = X::ld_nt(dz0, ob);
        o.dxm = X::ld_nt(dx1, ob); o.dym = X::ld_nt(dy1, ob); o.dzm = X::ld_nt(dz1, ob);
        o.exm = X::ld_nt(ex1, ob); o.eym = X::ld_nt(ey1, ob); o.ezm = X::ld_nt(ez1, ob);
        o.id = __builtin_nontemporal_load(ids + ob / ES);
        o.hh0 = 0.f; o.hh1 = 0.f;
        if (w_rlo) { const uint32_t oh = (uint32_t)k * szb + c_rlo; o.hh0 = X::ld(hxo, oh); o.hh1 = X::ld(hzo, oh); }
        if (w_col) { const uint32_t oh = (uint32_t)k * szb + c_col; o.hh0 = X::ld(hyo, oh); o.hh1 = X::ld(hzo, oh); }
    };

    // Step n -> n+1 starts one plane below the chunk (the second step's
    // D^{n+2} at k0 needs H^{n+3/2}, hence E^{n+1}, at k0-1) and ends one
    // above it (H^{n+3/2} at k1-1 needs E^{n+1} at k1).
    const int kAs = max(k0 - 1, 0);
    float hxm = 0.0f, hym = 0.0f, hzm = 0.0f;   // H^{n+1/2} of the own cell, plane kA-1
    int cur = 0;
    if (kAs > 0) {
        fetch_e(kAs - 1);
        stage_e(0);
        fetch_e(kAs);
        stage_e(1);
        __syncthreads();
        const int k = kAs - 1;
        const uint32_t ob = (uint32_t)k * szb + c_own;
        const float hx0 = X::ld(hxo, ob), hy0 = X::ld(hyo, ob);
        const float exk = sE[0][0][ty + 1][tx + 1], eyk = sE[0][1][ty + 1][tx + 1], ezk = sE[0][2][ty + 1][tx + 1];
        const float exk1 = sE[1][0][ty + 1][tx + 1], eyk1 = sE[1][1][ty + 1][tx + 1];
        const float ux = hx0 - chy * (sE[0][2][ty + 2][tx + 1] - ezk) + chy * (eyk1 - eyk);
        const float uy = hy0 - chz * (exk1 - exk) + chx * (sE[0][2][ty + 1][tx + 2] - ezk);
        hxm = (i >= 1 && i < nx - 1 && j < ny - 1 && k < nz - 1) ? ux : hx0;
        hym = (i < nx - 1 && j >= 1 && j < ny - 1 && k < nz - 1) ? uy : hy0;
        cur = 1;
        __syncthreads();
    } else {
        fetch_e(kAs);
        stage_e(0);
    }
    Own cu, nx_;
    fetch_own(kAs, cu);
    fetch_e(kAs + 1);

    const bool hx_ij = (i >= 1 && i < nx - 1 && j < ny - 1);
    const bool hy_ij = (i < nx - 1 && j >= 1 && j < ny - 1);
    const bool hz_ij = (i < nx - 1 && j < ny - 1);
    const bool dx_ij = (i < nx - 1 && j >= 1 && j < ny - 1);
    const bool dy_ij = (i >= 1 && i < nx - 1 && j < ny - 1);
    const bool dz_ij = (i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1);
    const bool e_i = (i >= 1);
    const int jh = j0 - 1;
    const bool rx_ij = (jh >= 0) && (i >= 1 && i < nx - 1 && jh < ny - 1);
    const bool rz_ij = (jh >= 0) && (i < nx - 1 && jh < ny - 1);
    const int ccr = min(max(cr, 1), R);
    const int jc = j0 - 1 + ccr;
    const bool col_ok = (side == 0) && (cr >= 1) && (cr <= R) && (i0 >= 1);
    const bool cy_ij = col_ok && (i0 - 1 < nx - 1) && (jc >= 1 && jc < ny - 1);
    const bool cz_ij = col_ok && (i0 - 1 < nx - 1) && (jc < ny - 1);
    float *col_dst_y = col_ok ? &sH[1][ccr][0] : &sJunk[0][tx];
    float *col_dst_z = col_ok ? &sH[2][ccr][0] : &sJunk[1][tx];

    // Mur faces inside the box: x = 0 at lane 1 of the first column of
    // tiles, x = nx-1 wherever it falls, y = 0 and ny-1 on whole waves.
    const bool xface_t = (i0 < 0) || (i0 <= nx - 1 && nx - 1 < i0 + W);
    const bool yface = (j == 0 || j == ny - 1);
    const float mdt = S.dt, mhy = S.hy, mhz = S.hz, mmu = S.mu0, mep = S.eps0;
    const long long nxy = (long long)nx * ny;

    float gxm = 0.0f, gym = 0.0f;              // H^{n+3/2} of the own cell, plane kA-2
    float pex = 0.f, pey = 0.f, pez = 0.f;     // E^n, own cell, plane kA-1
    float pdx = 0.f, pdy = 0.f, pdz = 0.f;     // D^n, own cell, plane kA-1
    float qdx = 0.f, qdy = 0.f, qdz = 0.f;     // D^{n+1} as its slot holds it, own cell, plane kA-1
    uint32_t pid = 0;                          // id, own cell, plane kA-1
    const float jzA = A.jz, jzB = S.jz2;

    for (int kA = kAs; kA <= k1; ++kA) {
        const bool realA = kA < nz;            // plane nz: only the second step's last plane is left
        const int nxt = cur ^ 1;
        stage_e(nxt);                          // E^n of plane kA+1
        const bool more = kA + 1 <= k1 && kA + 1 < nz;
        if (more) {
            fetch_own(kA + 1, nx_);
            fetch_e(kA + 2);
        }
        __syncthreads();  // E^n of planes kA and kA+1 staged

        // ---- step n -> n+1, plane kA: H^{n+1/2} of the box and its -x/-y halo
        const bool kh = (kA < nz - 1), kz = (kA >= 1 && kA < nz - 1);
        const float exk = sE[cur][0][ty + 1][tx + 1], eyk = sE[cur][1][ty + 1][tx + 1], ezk = sE[cur][2][ty + 1][tx + 1];
        float hx = 0.f, hy = 0.f, hz = 0.f;
        if (realA) {
            const float exk1 = sE[nxt][0][ty + 1][tx + 1], eyk1 = sE[nxt][1][ty + 1][tx + 1];
            const float ux = cu.hxo - chy * (sE[cur][2][ty + 2][tx + 1] - ezk) + chy * (eyk1 - eyk);
            const float uy = cu.hyo - chz * (exk1 - exk) + chx * (sE[cur][2][ty + 1][tx + 2] - ezk);
            const float uz = cu.hzo - chx * (sE[cur][1][ty + 1][tx + 2] - eyk) + chy * (sE[cur][0][ty + 2][tx + 1] - exk);
            hx = (hx_ij && kh) ? ux : cu.hxo;
            hy = (hy_ij && kh) ? uy : cu.hyo;
            hz = (hz_ij && kz) ? uz : cu.hzo;
            sH[0][ty + 1][tx + 1] = hx;
            sH[1][ty + 1][tx + 1] = hy;
            sH[2][ty + 1][tx + 1] = hz;
            if (w_rlo) {  // (i, j0-1): Hx, Hz
                const float ezh = sE[cur][2][0][tx + 1], eyh = sE[cur][1][0][tx + 1], exh = sE[cur][0][0][tx + 1];
                const float vx = cu.hh0 - chy * (sE[cur][2][1][tx + 1] - ezh) + chy * (sE[nxt][1][0][tx + 1] - eyh);
                const float vz = cu.hh1 - chx * (sE[cur][1][0][tx + 2] - eyh) + chy * (sE[cur][0][1][tx + 1] - exh);
                sH[0][0][tx + 1] = (rx_ij && kh) ? vx : cu.hh0;
                sH[2][0][tx + 1] = (rz_ij && kz) ? vz : cu.hh1;
            }
            if (w_col) {  // (i0-1, j0-1+ccr): Hy, Hz
                const float exh = sE[cur][0][ccr][0], eyh = sE[cur][1][ccr][0], ezh = sE[cur][2][ccr][0];
                const float vy = cu.hh0 - chz * (sE[nxt][0][ccr][0] - exh) + chx * (sE[cur][2][ccr][1] - ezh);
                const float vz = cu.hh1 - chx * (sE[cur][1][ccr][1] - eyh) + chy * (sE[cur][0][ccr + 1][0] - exh);
                *col_dst_y = (cy_ij && kh) ? vy : cu.hh0;
                *col_dst_z = (cz_ij && kz) ? vz : cu.hh1;
            }
        }
        __syncthreads();  // H^{n+1/2} of plane kA and its halo staged

        // ---- step n -> n+1, plane kA: D^{n+1} and E^{n+1} of the box, then
        // the x and z Mur faces (the y faces follow the next barrier)
        const uint32_t ob = (uint32_t)kA * szb + c_own;
        float fx = 0.f, fy = 0.f, fz = 0.f;         // E^{n+1}, own cell
        float dxu = 0.f, dyp = 0.f, dzp = 0.f;      // D^{n+1} as its slot holds it
        bool rdx = false, rdy = false, rdz = false, stx = false, sty = false, stz = false;
        if (realA) {
            rdx = dx_ij && kz;
            rdy = dy_ij && kz;
            rdz = (dz_ij && kh) || (cu.id & kSourceBit);
            const bool reyz = e_i && (kA >= 1);
            dxu = cu.dxn + cdy * (hz - sH[2][ty][tx + 1]) - cdz * (hy - hym);
            const float dyu = cu.dyn + cdz * (hx - hxm) - cdx * (hz - sH[2][ty + 1][tx]);
            float dzu = cu.dzn + cdx * (hy - sH[1][ty + 1][tx]) - cdy * (hx - sH[0][ty][tx + 1]);
            dzu = (cu.id & kSourceBit) ? dzu - jzA : dzu;
            dyp = dyu;
            dzp = dzu;
            // E cells whose D is not updated read the target slot's value.
            if (in && reyz && !rdy) dyp = X::ld(dyT, ob);
            if (in && reyz && !rdz) dzp = X::ld(dzT, ob);
            const int m = cu.id & kMaterialMask;
            const float ca = lA[m], cc = lC[m], cd = lD[m];
            const double g1 = lG1[m], g2 = lG2[m], g3 = lG3[m];
            fx = e_update(cc, cd, ca, g1, g2, g3, exk, cu.exm, dxu, cu.dxn, cu.dxm);
            fy = e_update(cc, cd, ca, g1, g2, g3, eyk, cu.eym, dyp, cu.dyn, cu.dym);
            fz = e_update(cc, cd, ca, g1, g2, g3, ezk, cu.ezm, dzp, cu.dzn, cu.dzm);
            // Cells outside the E update keep the target slot's value.
            if (in && !rdx) fx = X::ld(exT, ob);
            if (in && !reyz) { fy = X::ld(eyT, ob); fz = X::ld(ezT, ob); }
            stx = rdx;
            sty = reyz;
            stz = reyz;
            // Mur faces read E^{n+1} of their neighbour, a cell no face writes
            // (so its value before the faces is final), and E^n of both.
            if (xface_t) {
                const float eyr = __shfl_down(fy, 1, 64), ezr = __shfl_down(fz, 1, 64), ezl = __shfl_up(fz, 1, 64);
                if (i == 0 && j >= 0 && j < ny) {
                    const float4 c = A.murx0[(long long)kA * ny + j];
                    if (j < ny - 1 && kA >= 1 && kA < nz - 1) {   // Ey, x = 0
                        fy = c.x * ((c.y * eyr + c.z * sE[cur][1][ty + 1][tx + 2]) - c.w * eyk);
                        sty = true;
                    }
                    if (j >= 1 && j < ny - 1 && kA < nz - 1) {    // Ez, x = 0
                        fz = c.x * ((c.y * ezr + c.z * sE[cur][2][ty + 1][tx + 2]) - c.w * ezk);
                        stz = true;
                    }
                }
                if (i == nx - 1 && j >= 0 && j < ny) {
                    const float4 c = A.murx1[(long long)kA * ny + j];
                    if (j < ny - 1 && kA >= 1 && kA < nz - 1) {   // Ey, x = nx-1: the source neighbour twice
                        const float sn = sE[cur][1][ty + 1][tx];
                        fy = c.x * ((c.y * sn + c.z * sn) - c.w * eyk);
                        sty = true;
                    }
                    if (j >= 1 && j < ny - 1 && kA < nz - 1) {    // Ez, x = nx-1
                        fz = c.x * ((c.y * ezl + c.z * sE[cur][2][ty + 1][tx]) - c.w * ezk);
                        stz = true;
                    }
                }
            }
            if (kA == nz - 1 && in) {   // z = nz-1: neighbour E^{n+1} from the previous plane
                const long long p = (long long)j * nx + i;
                const int pb = (kA - 1) & 1;
                if (i < nx - 1 && j >= 1 && j < ny - 1) {
                    fx = mur1(mdt, mhz, mmu, mep, S.rpzEnd[nxy + p], S.rpzEnd[p], sF[pb][0][ty][tx], pex, exk);
                    stx = true;
                }
                if (i >= 1 && i < nx - 1 && j < ny - 1) {
                    fy = mur1(mdt, mhz, mmu, mep, S.rpzEnd[nxy + p], S.rpzEnd[p], sF[pb][1][ty][tx], pey, eyk);
                    sty = true;
                }
            }
            if (kA == 1 && kAs == 0 && in) {   // z = 0, one plane late: its neighbour is this plane
                const long long p = (long long)j * nx + i;
                if (i < nx - 1 && j >= 1 && j < ny - 1) {
                    const float v = mur1(mdt, mhz, mmu, mep, S.rpz0[p], S.rpz0[nxy + p], fx, exk, pex);
                    sF[0][0][ty][tx] = v;
                    if (own && k0 == 0) X::st_nt(exT, c_own, v);
                }
                if (i >= 1 && i < nx - 1 && j < ny - 1) {
                    const float v = mur1(mdt, mhz, mmu, mep, S.rpz0[p], S.rpz0[nxy + p], fy, eyk, pey);
                    sF[0][1][ty][tx] = v;
                    if (own && k0 == 0) X::st_nt(eyT, c_own, v);
                }
            }
            const int fb = kA & 1;
            sF[fb][0][ty][tx] = fx;
            sF[fb][1][ty][tx] = fy;
            sF[fb][2][ty][tx] = fz;
        }
        __syncthreads();  // E^{n+1} of plane kA staged (before its y faces)

        if (realA && yface) {   // whole waves on y = 0 or ny-1: neighbour row from LDS
            const int fb = kA & 1;
            const bool lo = (j == 0);
            const int jn = lo ? ty + 1 : ty - 1;
            const float *rp = lo ? S.rpy0 : S.rpyEnd;
            const long long py = (long long)kA * 2 * nx + i;
            const long long pf = py + (lo ? 0 : nx), pn = py + (lo ? nx : 0);
            if (i >= 0 && i < nx - 1 && kA >= 1 && kA < nz - 1) {   // Ex
                fx = mur1(mdt, mhy, mmu, mep, rp[pf], rp[pn], sF[fb][0][jn][tx], sE[cur][0][jn + 1][tx + 1], exk);
                sF[fb][0][ty][tx] = fx;
                stx = true;
            }
            if (i >= 1 && i < nx - 1 && kA < nz - 1) {              // Ez
                fz = mur1(mdt, mhy, mmu, mep, rp[pf], rp[pn], sF[fb][2][jn][tx], sE[cur][2][jn + 1][tx + 1], ezk);
                sF[fb][2][ty][tx] = fz;
                stz = true;
            }
        }
        if (realA && own && kA >= k0 && kA < k1) {
            if (rdx) X::st_nt(dxT, ob, dxu);
            if (rdy) X::st_nt(dyT, ob, dyp);
            if (rdz) X::st_nt(dzT, ob, dzp);
            if (stx) X::st_nt(exT, ob, fx);
            if (sty) X::st_nt(eyT, ob, fy);
            if (stz) X::st_nt(ezT, ob, fz);
        }

        // ---- step n+1 -> n+2, plane kB = kA-1: H^{n+3/2} of the box (its last
        // row and column are not needed and read padding)
        const int kB = kA - 1;
        const int gb1 = kB & 1;
        const float gxk = sF[gb1][0][ty][tx], gyk = sF[gb1][1][ty][tx], gzk = sF[gb1][2][ty][tx];
        float gx = 0.f, gy = 0.f, gz = 0.f;
        const bool khB = kB < nz - 1, kzB = kB >= 1 && kB < nz - 1;
        if (kB >= kAs) {
            const float vx = hxm - chy * (sF[gb1][2][ty + 1][tx] - gzk) + chy * (fy - gyk);
            const float vy = hym - chz * (fx - gxk) + chx * (sF[gb1][2][ty][tx + 1] - gzk);
            const float vz = hzm - chx * (sF[gb1][1][ty][tx + 1] - gyk) + chy * (sF[gb1][0][ty + 1][tx] - gxk);
            gx = (hx_ij && khB) ? vx : hxm;
            gy = (hy_ij && khB) ? vy : hym;
            gz = (hz_ij && kzB) ? vz : hzm;
            sG[0][ty + 1][tx + 1] = gx;
            sG[1][ty + 1][tx + 1] = gy;
            sG[2][ty + 1][tx + 1] = gz;
        }
        __syncthreads();  // H^{n+3/2} of plane kB staged

        // ---- step n+1 -> n+2, plane kB: D^{n+2}, E^{n+2} of the stored cells
        if (kB >= k0 && kB < k1) {
            const uint32_t obB = (uint32_t)kB * szb + c_own;
            const bool src = (pid & kSourceBit) != 0;
            const bool rdx2 = dx_ij && kzB, rdy2 = dy_ij && kzB, rdz2 = (dz_ij && khB) || src;
            const bool reyz2 = e_i && (kB >= 1);
            const float dxu2 = qdx + cdy * (gz - sG[2][ty][tx + 1]) - cdz * (gy - gym);
            const float dyu2 = qdy + cdz * (gx - gxm) - cdx * (gz - sG[2][ty + 1][tx]);
            float dzu2 = qdz + cdx * (gy - sG[1][ty + 1][tx]) - cdy * (gx - sG[0][ty][tx + 1]);
            dzu2 = src ? dzu2 - jzB : dzu2;
            float dyp2 = dyu2, dzp2 = dzu2;
            if (own && reyz2 && !rdy2) dyp2 = X::ld(dyS, obB);
            if (own && reyz2 && !rdz2) dzp2 = X::ld(dzS, obB);
            const int m = pid & kMaterialMask;
            const float ca = lA[m], cc = lC[m], cd = lD[m];
            const double g1 = lG1[m], g2 = lG2[m], g3 = lG3[m];
            const float exn = e_update(cc, cd, ca, g1, g2, g3, gxk, pex, dxu2, qdx, pdx);
            const float eyn = e_update(cc, cd, ca, g1, g2, g3, gyk, pey, dyp2, qdy, pdy);
            const float ezn = e_update(cc, cd, ca, g1, g2, g3, gzk, pez, dzp2, qdz, pdz);
            if (own) {
                X::st_nt(hxn, obB, gx); X::st_nt(hyn, obB, gy); X::st_nt(hzn, obB, gz);
                if (rdx2) { X::st_nt(dxS, obB, dxu2); X::st_nt(exS, obB, exn); }
                if (rdy2) X::st_nt(dyS, obB, dyu2);
                if (rdz2) X::st_nt(dzS, obB, dzu2);
                if (reyz2) { X::st_nt(eyS, obB, eyn); X::st_nt(ezS, obB, ezn); }
            }
        }
        if (kB >= kAs) { gxm = gx; gym = gy; }
        hxm = hx; hym = hy; hzm = hz;
        pex = exk; pey = eyk; pez = ezk;
        pdx = cu.dxn; pdy = cu.dyn; pdz = cu.dzn;
        qdx = dxu; qdy = dyp; qdz = dzp;
        pid = cu.id;
        cur = nxt;
        if (more) cu = nx_;
    }
}

// Two steps of one grid per launch: block b runs tile xcd_logical(b) at
// four waves per SIMD.
template <int R>
__global__ __launch_bounds__(64 * R) __attribute__((amdgpu_waves_per_eu(4))) void fused_step2(const Step2Args S)
{
    const StepArgs &A = S.A;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int nblk = A.tiles_x * A.tiles_y * A.tiles_z;
    const int b = xcd_logical((int)blockIdx.x, nblk, A.xcd_remap, xcd_unit(A.xcd_remap, A.tiles_x, A.tiles_y));
    const int bx = b % A.tiles_x;
    const int by = (b / A.tiles_x) % A.tiles_y;
    const int bz = b / (A.tiles_x * A.tiles_y);
    stage_lut<true>(A.lut, A.nmat, tid, 64 * R);
    fused2_tile<R>(S, bx, by, bz);
}

// Batch tables are read through the constant address space: they do not
// change while a launch runs, so the block-uniform reads become scalar
// loads and the grid's pointers stay in SGPRs.
template <typename T> using Const = const __attribute__((address_space(4))) T;
template <typename T> __device__ __forceinline__ Const<T> *as_const(const T *p) { return (Const<T> *)p; }

// One grid, one step, 16-bit storage with two cells per lane.
template <int TY, typename T>
__global__ __launch_bounds__(TX *TY) void fused_step_pair(const StepArgs A)
{
    const int tid = threadIdx.y * TX + threadIdx.x;
    const int nblk = A.tiles_x * A.tiles_y * A.tiles_z;
    const int b = xcd_logical((int)blockIdx.x, nblk, A.xcd_remap, xcd_unit(A.xcd_remap, A.tiles_x, A.tiles_y));
    const int bx = b % A.tiles_x;
    const int by = (b / A.tiles_x) % A.tiles_y;
    const int bz = b / (A.tiles_x * A.tiles_y);
    stage_lut<false>(A.lut, A.nmat, tid, TX * TY);
    fused_tile_pair<TY, T>(A, bx, by, bz);
}

// Step arguments of batch instance d at its step n: the ring slot and H
// buffer roles follow n exactly as a single grid's do after n steps.
__device__ __forceinline__ void batch_step_args(Const<BatchInst> &d, int n, StepArgs &A)
{
    const int t = n % 3, s1 = (n + 2) % 3, s2 = (n + 1) % 3, ho = n & 1, hn = ho ^ 1;
    A.hxo = d.H[ho][0]; A.hyo = d.H[ho][1]; A.hzo = d.H[ho][2];
    A.hxn = d.H[hn][0]; A.hyn = d.H[hn][1]; A.hzn = d.H[hn][2];
    A.ex0 = d.E[s1][0]; A.ey0 = d.E[s1][1]; A.ez0 = d.E[s1][2];
    A.ex1 = d.E[s2][0]; A.ey1 = d.E[s2][1]; A.ez1 = d.E[s2][2];
    A.exT = d.E[t][0]; A.eyT = d.E[t][1]; A.ezT = d.E[t][2];
    A.dx0 = d.D[s1][0]; A.dy0 = d.D[s1][1]; A.dz0 = d.D[s1][2];
    A.dx1 = d.D[s2][0]; A.dy1 = d.D[s2][1]; A.dz1 = d.D[s2][2];
    A.dxT = d.D[t][0]; A.dyT = d.D[t][1]; A.dzT = d.D[t][2];
    A.ids = d.ids;
    A.lut = d.lut;
    A.nmat = d.nmat;
    const int idx = d.jz_hold3 ? (n / 3) * 3 : n;
    A.jz = idx < d.jz_len ? d.jz[idx] : 0.0f;
    A.nx = d.nx; A.ny = d.ny; A.nz = d.nz;
    A.pitch = d.pitch;
    A.plane = d.plane;
    A.chx = d.ch[0]; A.chy = d.ch[1]; A.chz = d.ch[2];
    A.cdx = d.cd[0]; A.cdy = d.cd[1]; A.cdz = d.cd[2];
    A.kchunk = d.kchunk;
    A.tiles_x = d.tiles_x; A.tiles_y = d.tiles_y; A.tiles_z = d.tiles_z;
    A.xcd_remap = 0;
    A.guard = d.guard;
    A.mur_x = d.mur_x;
    A.murx0 = d.murx0;
    A.murx1 = d.murx1;
    A.zskip = d.zskip;
}

// Persistent fused kernel over every live instance of one storage class:
// block b runs work items b, b + gridDim.x, ... of the class's tile list
// (instances in list order, each in z-chunk-major tile order).  Tiles of
// instances outside their step window are skipped.
// The paired 16-bit form fits four waves per SIMD in 128 VGPRs without
// spilling (left alone it takes 142 and three).  The scaled-unit one-cell
// forms run five waves per SIMD in 96 VGPRs with a small spill (0-12 bytes
// per lane), measured 5% faster on 64 uneven fp16 grids than four waves;
// the exact fp32 form gains nothing from it and keeps four without spilling.
template <int TY, typename T, bool EXACT, bool PAIR>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(PAIR ? 4 : (EXACT ? 1 : 5)))) void fused_batch(
    const BatchLaunch L)
{
    const int tid = threadIdx.y * TX + threadIdx.x;
    Const<BatchClass> *C = as_const(L.cls);
    Const<BatchInst> *I = as_const(L.inst);
    const int total = C->total;
    const int step = as_const(L.sh)->step + L.off;
    int q = 0, staged = -1;
    const int G = (int)gridDim.x, b = (int)blockIdx.x;
    // Tile `it` of the loop goes to block it % G, on XCD it % 8 (G a
    // multiple of 8).  With runs of U tiles, XCD x takes run x of each group
    // of 8 runs, so U consecutive tiles of the class's list share an L2; a
    // partial last group keeps the identity.
    const int U = L.xcd_runs;
    const bool runs = U > 0 && G % 8 == 0;
    const int full = runs ? total / (8 * U) * (8 * U) : 0;
    for (int it = b; it < total; it += G) {
        const int t = (runs && it < full) ? it / (8 * U) * (8 * U) + (it % 8) * U + (it % (8 * U)) / 8 : it;
        while (C->start[q + 1] <= t) ++q;
        const int id = C->idx[q];
        Const<BatchInst> &d = I[id];
        const int n = step - d.start;
        if (n < 0 || n >= d.nsteps) continue;
        const int local = t - C->start[q];
        const int layer = d.tiles_x * d.tiles_y;
        const int bz = local / layer, r = local - bz * layer;
        const int by = r / d.tiles_x, bx = r - by * d.tiles_x;
        StepArgs A;
        batch_step_args(d, n, A);
        __syncthreads();  // LDS of the previous tile (and its table) no longer read
        if (id != staged) {
            stage_lut<EXACT>(A.lut, A.nmat, tid, TX * TY);
            staged = id;
        }
        if constexpr (PAIR) fused_tile_pair<TY, T>(A, bx, by, bz);
        else fused_tile<TY, T, EXACT, 64, EXACT>(A, bx, by, bz);   // zero-term rows per grid (A.zskip)
    }
}

// New value of face cell (i, j, k) of Mur segment s (fdtd_calculations.cu:
// 237-427): E'_f = mur1(rp_face, rp_nb, E'_nb, E^n_nb, E^n_f) with the
// neighbour one cell inward along the face normal.  Segments: 0, 1 Ex at
// j = 0, ny-1; 2, 3 Ex at k = 0, nz-1; 4, 5 Ey at i = 0, nx-1; 6, 7 Ey at
// k = 0, nz-1; 8, 9 Ez at i = 0, nx-1; 10, 11 Ez at j = 0, ny-1.  Every
// E'_nb read lands on a cell no segment writes.
template <typename T>
__device__ __forceinline__ float mur_value(const MurArgs &M, int s, int i, int j, int k)
{
    using X = Elem<T>;
    const T *exT = (const T *)M.exT, *eyT = (const T *)M.eyT, *ezT = (const T *)M.ezT;
    const T *exS = (const T *)M.exS, *eyS = (const T *)M.eyS, *ezS = (const T *)M.ezS;
    const int nx = M.nx, ny = M.ny;
    const long long sy = M.pitch, sz = M.plane;
    const float dt = M.dt, m0 = M.mu0, e0 = M.eps0;
    const long long o = (long long)k * sz + (long long)j * sy + i;
    auto G = [&](const T *p, long long q) { return X::ld(p, (uint32_t)(q * (long long)sizeof(T))); };
    auto rpy = [&](const float *p, int jj) { return p[(long long)k * 2 * nx + (long long)jj * nx + i]; };
    auto rpz = [&](const float *p, int kk) { return p[(long long)kk * ny * nx + (long long)j * nx + i]; };
    auto rpx = [&](const float *p, int ii) { return p[(long long)k * ny * 2 + (long long)j * 2 + ii]; };
    switch (s) {
    case 0: return mur1(dt, M.hy, m0, e0, rpy(M.rpy0, 0), rpy(M.rpy0, 1), G(exT, o + sy), G(exS, o + sy), G(exS, o));
    case 1: return mur1(dt, M.hy, m0, e0, rpy(M.rpyEnd, 1), rpy(M.rpyEnd, 0), G(exT, o - sy), G(exS, o - sy), G(exS, o));
    case 2: return mur1(dt, M.hz, m0, e0, rpz(M.rpz0, 0), rpz(M.rpz0, 1), G(exT, o + sz), G(exS, o + sz), G(exS, o));
    case 3: return mur1(dt, M.hz, m0, e0, rpz(M.rpzEnd, 1), rpz(M.rpzEnd, 0), G(exT, o - sz), G(exS, o - sz), G(exS, o));
    case 4: return mur1(dt, M.hx, m0, e0, rpx(M.rpx0, 0), rpx(M.rpx0, 1), G(eyT, o + 1), G(eyS, o + 1), G(eyS, o));
    // the x-high Ey face reads the source neighbour twice (fdtd_calculations.cu:323-325)
    case 5: return mur1(dt, M.hx, m0, e0, rpx(M.rpxEnd, 1), rpx(M.rpxEnd, 0), G(eyS, o - 1), G(eyS, o - 1), G(eyS, o));
    case 6: return mur1(dt, M.hz, m0, e0, rpz(M.rpz0, 0), rpz(M.rpz0, 1), G(eyT, o + sz), G(eyS, o + sz), G(eyS, o));
    case 7: return mur1(dt, M.hz, m0, e0, rpz(M.rpzEnd, 1), rpz(M.rpzEnd, 0), G(eyT, o - sz), G(eyS, o - sz), G(eyS, o));
    case 8: return mur1(dt, M.hx, m0, e0, rpx(M.rpx0, 0), rpx(M.rpx0, 1), G(ezT, o + 1), G(ezS, o + 1), G(ezS, o));
    case 9: return mur1(dt, M.hx, m0, e0, rpx(M.rpxEnd, 1), rpx(M.rpxEnd, 0), G(ezT, o - 1), G(ezS, o - 1), G(ezS, o));
    case 10: return mur1(dt, M.hy, m0, e0, rpy(M.rpy0, 0), rpy(M.rpy0, 1), G(ezT, o + sy), G(ezS, o + sy), G(ezS, o));
    default: return mur1(dt, M.hy, m0, e0, rpy(M.rpyEnd, 1), rpy(M.rpyEnd, 0), G(ezT, o - sy), G(ezS, o - sy), G(ezS, o));
    }
}

// The Mur segment that holds cell (i, j, k) of E component c (0 Ex, 1 Ey,
// 2 Ez), or -1.  The ranges are the reference's (fdtd_calculations.cu:
// 237-427); for each component they are disjoint.
__device__ __forceinline__ int mur_segment_of(int c, int i, int j, int k, int nx, int ny, int nz)
{
    if (c == 0) {
        if (i >= nx - 1) return -1;
        if ((j == 0 || j == ny - 1) && k >= 1 && k < nz - 1) return j == 0 ? 0 : 1;
        if ((k == 0 || k == nz - 1) && j >= 1 && j < ny - 1) return k == 0 ? 2 : 3;
        return -1;
    }
    if (c == 1) {
        if ((i == 0 || i == nx - 1) && j < ny - 1 && k >= 1 && k < nz - 1) return i == 0 ? 4 : 5;
        if ((k == 0 || k == nz - 1) && i >= 1 && i < nx - 1 && j < ny - 1) return k == 0 ? 6 : 7;
        return -1;
    }
    if ((i == 0 || i == nx - 1) && j >= 1 && j < ny - 1 && k < nz - 1) return i == 0 ? 8 : 9;
    if ((j == 0 || j == ny - 1) && i >= 1 && i < nx - 1 && k < nz - 1) return j == 0 ? 10 : 11;
    return -1;
}

// Face cell t (in segment order) of the 12 (component, face) Mur updates;
// seg holds the 13 segment prefix sums (M.seg_start is not read).
template <typename T, typename Seg>
__device__ __forceinline__ void mur_cell(const MurArgs &M, Seg seg, int t)
{
    if (t >= seg[12]) return;
    // the last segment starting at or before t (prefix sums never decrease;
    // no dynamic index into seg, which would copy the arguments to scratch)
    int s = 0, base = 0;
#pragma unroll
    for (int q = 1; q < 12; q++)
        if (t >= seg[q]) { s = q; base = seg[q]; }
    const int q = t - base;
    const int nx = M.nx, ny = M.ny, nz = M.nz;
    int i, j, k;
    switch (s) {
    case 0: case 1: i = q % (nx - 1); k = 1 + q / (nx - 1); j = s == 0 ? 0 : ny - 1; break;   // i in [0,nx-1), k in [1,nz-1)
    case 2: case 3: i = q % (nx - 1); j = 1 + q / (nx - 1); k = s == 2 ? 0 : nz - 1; break;   // i in [0,nx-1), j in [1,ny-1)
    case 4: case 5: j = q % (ny - 1); k = 1 + q / (ny - 1); i = s == 4 ? 0 : nx - 1; break;   // j in [0,ny-1), k in [1,nz-1)
    case 6: case 7: i = 1 + q % (nx - 2); j = q / (nx - 2); k = s == 6 ? 0 : nz - 1; break;   // i in [1,nx-1), j in [0,ny-1)
    case 8: case 9: j = 1 + q % (ny - 2); k = q / (ny - 2); i = s == 8 ? 0 : nx - 1; break;   // j in [1,ny-1), k in [0,nz-1)
    default: i = 1 + q % (nx - 2); k = q / (nx - 2); j = s == 10 ? 0 : ny - 1; break;         // i in [1,nx-1), k in [0,nz-1)
    }
    const float v = mur_value<T>(M, s, i, j, k);
    T *dst = (T *)(s < 4 ? M.exT : (s < 8 ? M.eyT : M.ezT));
    const long long o = (long long)k * M.plane + (long long)j * M.pitch + i;
    Elem<T>::st(dst, (uint32_t)(o * (long long)sizeof(T)), v);
}

// Probe value of field f (Dx Dy Dz Hx Hy Hz Ex Ey Ez) at (i, j, k) after this
// step's Mur faces, read in the same launch as the faces are written: the
// stored value, except for an E cell of a segment this launch updates, whose
// new value is evaluated here from the same operands (so it equals, bit for
// bit, what the face thread stores).
template <typename T, typename Seg>
__device__ __forceinline__ float probe_value(const MurArgs &M, Seg seg, const void *field, int f, int i, int j, int k)
{
    if (f >= 6) {
        const int s = mur_segment_of(f - 6, i, j, k, M.nx, M.ny, M.nz);
        bool active = false;
#pragma unroll
        for (int q = 0; q < 12; q++)
            if (q == s) active = seg[q + 1] > seg[q];
        if (active) return Elem<T>::rnd(mur_value<T>(M, s, i, j, k));   // as mur_cell stores it
    }
    const long long o = (long long)k * M.plane + (long long)j * M.pitch + i;
    return Elem<T>::ld((const T *)field, (uint32_t)(o * (long long)sizeof(T)));
}

template <typename T>
__global__ __launch_bounds__(kMurBlock) void mur_faces(const MurArgs M)
{
    mur_cell<T>(M, M.seg_start, blockIdx.x * kMurBlock + threadIdx.x);
}

// Mur faces, then (in the blocks after them) the step's probe row in the
// same launch: the probe threads do not wait for the face threads, since
// probe_value evaluates an updated face cell itself.
template <typename T>
__global__ __launch_bounds__(kMurBlock) void mur_probe(const MurArgs M, const ProbeArgs P)
{
    const int mb = (M.seg_start[12] + kMurBlock - 1) / kMurBlock;
    if ((int)blockIdx.x < mb) {
        mur_cell<T>(M, M.seg_start, blockIdx.x * kMurBlock + threadIdx.x);
        return;
    }
    const int p = ((int)blockIdx.x - mb) * kMurBlock + (int)threadIdx.x;
    if (p >= P.nprobe) return;
    const int *q = P.probes + 4 * p;
    const int f = q[0];
    const float v = probe_value<T>(M, M.seg_start, P.field[f], f, q[1], q[2], q[3]);
    P.out[(long long)P.row * P.nprobe + p] = (double)v * P.factor[f];
}

// Mur operands of batch instance d at its step n (ring slot n % 3 written,
// (n - 1) % 3 the source).
template <typename Inst>
__device__ __forceinline__ void batch_mur_args(const Inst &d, int n, MurArgs &M)
{
    const int tt = n % 3, s1 = (n + 2) % 3;
    M.exT = d.E[tt][0]; M.eyT = d.E[tt][1]; M.ezT = d.E[tt][2];
    M.exS = d.E[s1][0]; M.eyS = d.E[s1][1]; M.ezS = d.E[s1][2];
    M.rpx0 = d.rp[0]; M.rpy0 = d.rp[1]; M.rpz0 = d.rp[2];
    M.rpxEnd = d.rp[3]; M.rpyEnd = d.rp[4]; M.rpzEnd = d.rp[5];
    M.nx = d.nx; M.ny = d.ny; M.nz = d.nz; M.pitch = d.pitch; M.plane = d.plane;
    M.dt = d.dt; M.hx = d.hx; M.hy = d.hy; M.hz = d.hz; M.mu0 = d.mu0; M.eps0 = d.eps0;
}

// Mur faces and probe rows of every live instance of a batch.  Block b runs
// work items b, b + gridDim.x, ... of the shared Mur list (kMurBlock face
// cells each), then probe records r = (instance, probe) in grid-stride
// order; record r writes row n of the instance's series after its step n
// (probe_value: no wait for the face threads).
__global__ __launch_bounds__(kMurBlock) void mur_batch(const BatchLaunch L)
{
    Const<BatchShared> *S = as_const(L.sh);
    const int total = S->mur_total;
    const int step = S->step + L.off;
    int q = 0;
    for (int blk = blockIdx.x; blk < total; blk += gridDim.x) {
        while (S->mur_start[q + 1] <= blk) ++q;
        Const<BatchInst> &d = as_const(L.inst)[S->mur_idx[q]];
        const int n = step - d.start;
        if (n < 0 || n >= d.nsteps) continue;
        MurArgs M;
        batch_mur_args(d, n, M);
        Const<int> *seg = d.mur_seg;
        const int t = (blk - S->mur_start[q]) * kMurBlock + (int)threadIdx.x;
        switch (d.storage) {
        case kStoreF16: mur_cell<__half>(M, seg, t); break;
        case kStoreBF16: mur_cell<__hip_bfloat16>(M, seg, t); break;
        default: mur_cell<float>(M, seg, t); break;
        }
    }
    const int nrec = S->nrec;
    for (int r = blockIdx.x * kMurBlock + threadIdx.x; r < nrec; r += gridDim.x * kMurBlock) {
        const int id = L.sh->rec[r][0], p = L.sh->rec[r][1];
        const BatchInst &d = L.inst[id];
        const int n = step - d.start;
        if (n < 0 || n >= d.nsteps || n >= d.probe_rows) continue;
        MurArgs M;
        batch_mur_args(d, n, M);
        const int *pq = d.probes + 4 * p;
        const int f = pq[0];
        const void *base = f < 3 ? d.D[n % 3][f] : (f < 6 ? d.H[(n & 1) ^ 1][f - 3] : d.E[n % 3][f - 6]);
        float v;
        switch (d.storage) {
        case kStoreF16: v = probe_value<__half>(M, d.mur_seg, base, f, pq[1], pq[2], pq[3]); break;
        case kStoreBF16: v = probe_value<__hip_bfloat16>(M, d.mur_seg, base, f, pq[1], pq[2], pq[3]); break;
        default: v = probe_value<float>(M, d.mur_seg, base, f, pq[1], pq[2], pq[3]); break;
        }
        d.probe_out[(long long)n * d.nprobe + p] = (double)v * d.factor[f];
    }
}

__global__ void batch_advance(BatchShared *sh, int steps)
{
    if (threadIdx.x == 0) sh->step += steps;
}

__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Block-wide sum in a fixed order (lane tree within waves, then waves in
// order); the result is valid in thread 0.
template <int NT>
__device__ inline double block_sum(double v)
{
    __shared__ double part[NT / 64];
    v = wave_sum(v);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) part[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < NT / 64; q++) t += part[q];
    return t;
}

template <typename T>
__global__ __launch_bounds__(256) void energy_partial(const EnergyArgs A)
{
    using X = Elem<T>;
    const long long nxy = (long long)A.nx * A.ny, n = nxy * A.nz;
    double ed = 0.0, hh = 0.0;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
        const int k = (int)(q / nxy);
        const long long r = q - (long long)k * nxy;
        const int j = (int)(r / A.nx), i = (int)(r - (long long)j * A.nx);
        const uint32_t b = (uint32_t)(((long long)k * A.plane + (long long)j * A.pitch + i) * (long long)sizeof(T));
        for (int c = 0; c < 3; c++) {
            const double e = X::ld((const T *)A.E[c], b), d = X::ld((const T *)A.D[c], b), h = X::ld((const T *)A.H[c], b);
            ed += e * d;
            hh += h * h;
        }
    }
    const double t = block_sum<256>(0.5 * ed * A.ed_factor + 0.5 * hh * A.hh_factor);
    if (threadIdx.x == 0) A.partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void energy_final(const EnergyArgs A)
{
    double v = 0.0;
    for (int q = threadIdx.x; q < kEnergyBlocks; q += 256) v += A.partial[q];
    const double t = block_sum<256>(v);
    if (threadIdx.x == 0) *A.out = t;
}

template <typename T>
__global__ void probe_cells(const ProbeArgs P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.nprobe) return;
    const int *q = P.probes + 4 * p;
    const long long o = (long long)q[3] * P.plane + (long long)q[2] * P.pitch + q[1];
    const float v = Elem<T>::ld((const T *)P.field[q[0]], (uint32_t)(o * (long long)sizeof(T)));
    P.out[(long long)P.row * P.nprobe + p] = (double)v * P.factor[q[0]];
}

}  // namespace

void mur_segments(int nx, int ny, int nz, int *seg, unsigned skip)
{
    const int cnt[12] = {
        (nx - 1) * (nz - 2), (nx - 1) * (nz - 2), (nx - 1) * (ny - 2), (nx - 1) * (ny - 2),
        (ny - 1) * (nz - 2), (ny - 1) * (nz - 2), (nx - 2) * (ny - 1), (nx - 2) * (ny - 1),
        (ny - 2) * (nz - 1), (ny - 2) * (nz - 1), (nx - 2) * (nz - 1), (nx - 2) * (nz - 1),
    };
    seg[0] = 0;
    for (int q = 0; q < 12; q++) seg[q + 1] = seg[q] + ((skip >> q & 1u) ? 0 : std::max(cnt[q], 0));
}

hipError_t launch_mur(const MurArgs &a, int storage, hipStream_t s)
{
    const int total = a.seg_start[12];
    if (total <= 0) return hipSuccess;
    const int grid = (total + kMurBlock - 1) / kMurBlock;
    switch (storage) {
    case kStoreF32:
    case kStoreF32Fast: mur_faces<float><<<grid, kMurBlock, 0, s>>>(a); break;
    case kStoreF16: mur_faces<__half><<<grid, kMurBlock, 0, s>>>(a); break;
    case kStoreBF16: mur_faces<__hip_bfloat16><<<grid, kMurBlock, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

namespace {

// Copies up to kHaloFields equal-sized runs of 16-byte words, one run per
// blockIdx.y (halo planes into a message buffer and back).
__global__ __launch_bounds__(256) void plane_copy(const PlaneCopy P)
{
    const int f = (int)blockIdx.y;
    const uint4 *src = P.src[0];
    uint4 *dst = P.dst[0];
#pragma unroll
    for (int q = 1; q < kHaloFields; q++)
        if (f == q) { src = P.src[q]; dst = P.dst[q]; }
    for (long long w = (long long)blockIdx.x * 256 + threadIdx.x; w < P.words; w += (long long)gridDim.x * 256)
        dst[w] = src[w];
}

// Unpitches (and for scaled storage, unscales) the nine fields of a context
// into packed fp32 volumes: lanes run along x, blockIdx.y is the field.
template <typename T, bool SCALED>
__global__ __launch_bounds__(256) void field_snapshot(const SnapArgs A)
{
    const int f = (int)blockIdx.y;
    const T *src = (const T *)A.field[0];
    double factor = A.factor[0];
#pragma unroll
    for (int q = 1; q < 9; q++)
        if (f == q) { src = (const T *)A.field[q]; factor = A.factor[q]; }
    const long long n = (long long)A.nx * A.ny * A.nz;
    float *out = A.out + (long long)f * n;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
        const long long row = t / A.nx;
        const int i = (int)(t - row * A.nx);
        const long long at = row * A.pitch + i;   // row = k*ny + j, and plane = pitch*ny
        if constexpr (SCALED) {
            float v;
            if constexpr (std::is_same<T, float>::value) v = src[at];
            else v = Elem<T>::cvt(((const unsigned short *)src)[at]);
            out[t] = (float)((double)v * factor);
        } else {
            out[t] = src[at];
        }
    }
}

}  // namespace

hipError_t launch_snapshot(const SnapArgs &a, int storage, hipStream_t s)
{
    const long long n = (long long)a.nx * a.ny * a.nz;
    if (n <= 0 || !a.out) return hipErrorInvalidValue;
    const dim3 grid((unsigned)std::min<long long>((n + 255) / 256, 4096), 9);
    switch (storage) {
    case kStoreF32: field_snapshot<float, false><<<grid, 256, 0, s>>>(a); break;
    case kStoreF32Fast: field_snapshot<float, true><<<grid, 256, 0, s>>>(a); break;
    case kStoreF16: field_snapshot<__half, true><<<grid, 256, 0, s>>>(a); break;
    case kStoreBF16: field_snapshot<__hip_bfloat16, true><<<grid, 256, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_plane_copy(const PlaneCopy &p, hipStream_t s)
{
    if (p.n <= 0 || p.n > kHaloFields || p.words <= 0) return hipErrorInvalidValue;
    const long long blocks = std::min<long long>((p.words + 255) / 256, 1024);
    plane_copy<<<dim3((unsigned)blocks, (unsigned)p.n), 256, 0, s>>>(p);
    return hipGetLastError();
}

hipError_t launch_mur_probes(const MurArgs &a, const ProbeArgs &p, int storage, hipStream_t s)
{
    const int mb = (a.seg_start[12] + kMurBlock - 1) / kMurBlock;
    const int grid = mb + (p.nprobe + kMurBlock - 1) / kMurBlock;
    if (grid <= 0) return hipSuccess;
    switch (storage) {
    case kStoreF32:
    case kStoreF32Fast: mur_probe<float><<<grid, kMurBlock, 0, s>>>(a, p); break;
    case kStoreF16: mur_probe<__half><<<grid, kMurBlock, 0, s>>>(a, p); break;
    case kStoreBF16: mur_probe<__hip_bfloat16><<<grid, kMurBlock, 0, s>>>(a, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_energy(const EnergyArgs &a, int storage, hipStream_t s)
{
    switch (storage) {
    case kStoreF32:
    case kStoreF32Fast: energy_partial<float><<<kEnergyBlocks, 256, 0, s>>>(a); break;
    case kStoreF16: energy_partial<__half><<<kEnergyBlocks, 256, 0, s>>>(a); break;
    case kStoreBF16: energy_partial<__hip_bfloat16><<<kEnergyBlocks, 256, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    energy_final<<<1, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_probes(const ProbeArgs &a, int storage, hipStream_t s)
{
    if (a.nprobe <= 0) return hipSuccess;
    const int grid = (a.nprobe + 63) / 64;
    switch (storage) {
    case kStoreF32:
    case kStoreF32Fast: probe_cells<float><<<grid, 64, 0, s>>>(a); break;
    case kStoreF16: probe_cells<__half><<<grid, 64, 0, s>>>(a); break;
    case kStoreBF16: probe_cells<__hip_bfloat16><<<grid, 64, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

namespace {

// One fused launch.  With events, they take the dispatch's own start and
// end timestamps (no marker packets around it).
hipError_t fire(const void *kernel, unsigned nblk, dim3 block, size_t lds, hipStream_t s, const StepArgs &a,
                hipEvent_t t0, hipEvent_t t1)
{
    StepArgs arg = a;
    void *args[] = {&arg};
    return hipExtLaunchKernel(kernel, dim3(nblk), block, args, lds, s, t0, t1, 0);
}

template <typename T, bool EXACT>
hipError_t launch_fused_t(const StepArgs &a, int ty, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    const unsigned nblk = (unsigned)a.tiles_x * a.tiles_y * a.tiles_z;
    const size_t lds = (size_t)a.nmat * (3 * sizeof(double) + 3 * sizeof(float));
    auto go = [&](const void *k, int w, int rows) { return fire(k, nblk, dim3(w, rows), lds, s, a, t0, t1); };
    if constexpr (sizeof(T) == 2) {
        if (a.cpl == 2) {
            switch (ty) {
            case 2: return go((const void *)fused_step_pair<2, T>, TX, 2);
            case 4: return go((const void *)fused_step_pair<4, T>, TX, 4);
            case 8: return go((const void *)fused_step_pair<8, T>, TX, 8);
            default: return hipErrorInvalidValue;
            }
        }
    }
    if constexpr (EXACT) {
        if (a.zskip) {   // zero-term rows
            if (a.txw == 128) {
                switch (ty) {
                case 2: return go((const void *)fused_step<2, T, EXACT, 128, true>, 128, 2);
                case 4: return go((const void *)fused_step<4, T, EXACT, 128, true>, 128, 4);
                case 8: return go((const void *)fused_step<8, T, EXACT, 128, true>, 128, 8);
                default: return hipErrorInvalidValue;
                }
            }
            switch (ty) {
            case 2: return go((const void *)fused_step<2, T, EXACT, 64, true>, TX, 2);
            case 4: return go((const void *)fused_step<4, T, EXACT, 64, true>, TX, 4);
            case 8: return go((const void *)fused_step<8, T, EXACT, 64, true>, TX, 8);
            default: return hipErrorInvalidValue;
            }
        }
    }
    if (a.txw == 128) {
        switch (ty) {
        case 2: return go((const void *)fused_step<2, T, EXACT, 128>, 128, 2);
        case 4: return go((const void *)fused_step<4, T, EXACT, 128>, 128, 4);
        case 8: return go((const void *)fused_step<8, T, EXACT, 128>, 128, 8);
        default: return hipErrorInvalidValue;
        }
    }
    switch (ty) {
    case 2: return go((const void *)fused_step<2, T, EXACT, 64>, TX, 2);
    case 4: return go((const void *)fused_step<4, T, EXACT, 64>, TX, 4);
    case 8: return go((const void *)fused_step<8, T, EXACT, 64>, TX, 8);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_fused_step(const StepArgs &a, int ty, int storage, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    switch (storage) {
    case kStoreF32: return launch_fused_t<float, true>(a, ty, s, t0, t1);
    case kStoreF32Fast: return launch_fused_t<float, false>(a, ty, s, t0, t1);
    case kStoreF16: return launch_fused_t<__half, false>(a, ty, s, t0, t1);
    case kStoreBF16: return launch_fused_t<__hip_bfloat16, false>(a, ty, s, t0, t1);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fused_step2(const Step2Args &a, int rows, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    const unsigned nblk = (unsigned)a.A.tiles_x * a.A.tiles_y * a.A.tiles_z;
    const size_t lds = (size_t)a.A.nmat * (3 * sizeof(double) + 3 * sizeof(float));
    const void *k = nullptr;
    switch (rows) {
    case 6: k = (const void *)fused_step2<6>; break;
    case 8: k = (const void *)fused_step2<8>; break;
    case 10: k = (const void *)fused_step2<10>; break;
    default: return hipErrorInvalidValue;
    }
    Step2Args arg = a;
    void *args[] = {&arg};
    return hipExtLaunchKernel(k, dim3(nblk), dim3(64, rows), args, lds, s, t0, t1, 0);
}

namespace {

template <int TY, typename T, bool EXACT, bool PAIR>
hipError_t launch_fused_batch_t(const BatchLaunch &l, int grid, hipStream_t s)
{
    fused_batch<TY, T, EXACT, PAIR><<<grid, dim3(TX, TY), batch_fused_lds(), s>>>(l);
    return hipGetLastError();
}

template <typename T, bool EXACT, bool PAIR>
hipError_t launch_fused_batch_s(const BatchLaunch &l, int ty, int grid, hipStream_t s)
{
    switch (ty) {
    case 2: return launch_fused_batch_t<2, T, EXACT, PAIR>(l, grid, s);
    case 4: return launch_fused_batch_t<4, T, EXACT, PAIR>(l, grid, s);
    case 8: return launch_fused_batch_t<8, T, EXACT, PAIR>(l, grid, s);
    default: return hipErrorInvalidValue;
    }
}

template <int TY, typename T, bool EXACT, bool PAIR>
int occupancy_t()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fused_batch<TY, T, EXACT, PAIR>, TX * TY,
                                                     batch_fused_lds()) != hipSuccess)
        return 0;
    return nb;
}

template <typename T, bool EXACT, bool PAIR>
int occupancy_s(int ty)
{
    switch (ty) {
    case 2: return occupancy_t<2, T, EXACT, PAIR>();
    case 4: return occupancy_t<4, T, EXACT, PAIR>();
    case 8: return occupancy_t<8, T, EXACT, PAIR>();
    default: return 0;
    }
}

}  // namespace

size_t batch_fused_lds() { return (size_t)kMaxMaterials * (3 * sizeof(double) + 3 * sizeof(float)); }

hipError_t launch_fused_batch(const BatchLaunch &l, int ty, int cls, int grid, hipStream_t s)
{
    switch (cls) {
    case 0: return launch_fused_batch_s<float, true, false>(l, ty, grid, s);
    case 1: return launch_fused_batch_s<__half, false, false>(l, ty, grid, s);
    case 2: return launch_fused_batch_s<__hip_bfloat16, false, false>(l, ty, grid, s);
    case 3: return launch_fused_batch_s<float, false, false>(l, ty, grid, s);
    case 4: return launch_fused_batch_s<__half, false, true>(l, ty, grid, s);
    case 5: return launch_fused_batch_s<__hip_bfloat16, false, true>(l, ty, grid, s);
    default: return hipErrorInvalidValue;
    }
}

int batch_fused_occupancy(int ty, int cls)
{
    switch (cls) {
    case 0: return occupancy_s<float, true, false>(ty);
    case 1: return occupancy_s<__half, false, false>(ty);
    case 2: return occupancy_s<__hip_bfloat16, false, false>(ty);
    case 3: return occupancy_s<float, false, false>(ty);
    case 4: return occupancy_s<__half, false, true>(ty);
    case 5: return occupancy_s<__hip_bfloat16, false, true>(ty);
    default: return 0;
    }
}

hipError_t launch_mur_batch(const BatchLaunch &l, int grid, hipStream_t s)
{
    mur_batch<<<grid, kMurBlock, 0, s>>>(l);
    return hipGetLastError();
}


hipError_t launch_batch_advance(BatchShared *sh, int steps, hipStream_t s)
{
    batch_advance<<<1, 64, 0, s>>>(sh, steps);
    return hipGetLastError();
}

namespace {

// The fused step's memory stream without its arithmetic: 15 arrays read and
// 9 written, element for element, every access nontemporal; a grid-stride
// loop over n elements of each array.
// The same mix over row records: the 24 rows of `row` cells that one row of
// the grid touches are contiguous, so a wave's 24 accesses fall in one
// record instead of 24 arrays far apart.
__global__ __launch_bounds__(256) void stream_mix_rows(const StreamMix p)
{
    const float *base = p.r[0];
    float *wbase = const_cast<float *>(base);
    const long long W = p.row;
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < p.n; q += (long long)gridDim.x * 256) {
        const long long r = q / W, x = q - r * W;
        const long long rec = r * 24 * W + x;
        float acc = 0.0f;
#pragma unroll
        for (int a = 0; a < 15; a++) acc += __builtin_nontemporal_load(&base[rec + a * W]);
#pragma unroll
        for (int b = 15; b < 24; b++) __builtin_nontemporal_store(acc, &wbase[rec + b * W]);
    }
}

__global__ __launch_bounds__(256) void stream_mix(const StreamMix p)
{
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < p.n; q += (long long)gridDim.x * 256) {
        float acc = 0.0f;
#pragma unroll
        for (int a = 0; a < 15; a++) acc += __builtin_nontemporal_load(&p.r[a][q]);
#pragma unroll
        for (int b = 0; b < 9; b++) __builtin_nontemporal_store(acc, &p.w[b][q]);
    }
}

}  // namespace

hipError_t launch_stream_mix(const StreamMix &p, int blocks, hipStream_t s)
{
    if (p.row > 0) stream_mix_rows<<<blocks, 256, 0, s>>>(p);
    else stream_mix<<<blocks, 256, 0, s>>>(p);
    return hipGetLastError();
}

int stream_mix_resident_blocks()
{
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, stream_mix, 256, 0) != hipSuccess) return 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return nb * cus;
}
