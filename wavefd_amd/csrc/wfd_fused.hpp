// wfd_fused.hpp -- single-pass fused time step (D+ sandwich, D-, leapfrog with Q) for the
// three physics.  One kernel replaces the reference's two OpenMP passes
// (e.g. src/prop3DAcoTTIDenQ_DEO2_FDTD.h:359-379): the D+ temporaries never reach HBM.
//
// Decomposition ("2.5D", marching along x, the slowest axis):
//   * a workgroup of ZW x NROW threads owns a (y,z) tile of temporaries, ZW cells along z
//     (lanes, coalesced) by NROW rows along y; its outputs are the inner (ZW-7) x (NROW-7)
//     cells (the backward difference D- reads temporaries at offsets -4..+3), and it marches
//     over a chunk of x-planes;
//   * every thread keeps, in registers, p(x..x+7) of its column (the D+x operand) and the
//     x-temporaries of planes x-4..x+3 (the D-x operand);
//   * at step x the workgroup stages plane x+3 of p (and m) in LDS with a 7-cell border,
//     computes the full gradient and the material sandwich at (x+3, y, z) for the whole tile,
//     keeps the y/z temporaries in a 4-plane LDS ring, and the output threads then apply D-
//     and the leapfrog update at plane x.
// Halo threads recompute the sandwich of their neighbours' cells, which is what lets TTI's
// rotated sandwich (full gradient at every temporary) fuse at all.
//
// Memory: every global access is a uniform plane base plus a 32-bit byte offset inside the
// plane computed once per thread; out-of-grid planes and columns are clamped for the address
// and zeroed, so no access needs a bounds branch.  Every
// operand of step x+1 (p(x+8), the border of plane x+4, the material at x+4, v/q/pOld at x+1)
// is requested before step x's barrier, one step ahead of its use.  One barrier per step:
// the p-plane buffer is double-buffered and the temporaries live in a 4-plane ring.
// Workgroups are numbered so that each XCD (workgroup id mod 8) sweeps a contiguous run of
// tiles, which keeps neighbouring tiles, whose halos overlap, in the same L2.
//
// Semantics match the operator kernels exactly: temporaries are 0 outside the update
// region (the reference's zeroed annulus), the free surface uses the odd mirror for D+ and
// the even mirror for D-, x/y derivatives vanish at kz=0 and L=0 there.  Expression order
// is unchanged, so results are bit-identical to the two-pass path.
#pragma once
#include <type_traits>

#include "wfd_kernels.hpp"
#include "wfd_runtime.hpp"

namespace wfd {

template <int PHYS>
struct FusedArgs {
    Grid g;
    float dt2;
    int cx;        // x-planes per chunk: chunks [0, nlong) hold cx + cxu planes, the last one up to nx
    int nlong;
    int cxu;       // the kernel's x-loop unrolling (its steps come in multiples of cxu)
    int nzt, nyt;  // tiles along z and y
    int nxc;       // x-chunks
    int bz, by;    // tiles per block of the XCD-local order (bz = 0: plain z-fastest, -1: y-fastest)
    int per_xcd;     // workgroups per XCD that run at once (rounds kept in step); 0: no waiting
    int slack;       // a round starts when all but this many workgroups of the earlier rounds finished
    int xmap;        // 0: each XCD a contiguous run of tiles; G > 0: groups of G tiles dealt to the XCDs in turn
                     // (WAVEFD_FUSED_SLACK sets it, for measurements)
    unsigned *sync;  // per-XCD count of finished workgroups, zeroed before the launch
    const float *pCur, *mCur;
    float *pOld, *mOld, *pSpace, *mSpace;
    const float *v, *b, *q, *eps, *eta, *f, *sT, *cT, *sP, *cP;
    // Q by sponge distance (nsponge + 1 values) while q holds the SetupDtOmegaInvQ profile, else
    // null: the update then takes qtab[min(distance, nsponge)] instead of loading q
    const float *qtab;
    int nsponge;
    // epilogue operands (see Epilogue)
    float dt;
    float *e_dp, *e_dm;             // SAVE_DP: DP, DM out; IMAGE: g (e_dp) in/out
    const float *e_dv, *e_ip, *e_im;  // BORN_INJ: dV; BORN_INJ / IMAGE: DP, DM in
};

// Work fused after the update of each output cell (the Born loops' per-step calls, fused so
// that their operands are read once with the step's):
//   NONE      the reference's TimeStep: pOld <- p+, pSpace <- L (mSpace)
//   SAVE_DP   TimeStep then ScaleSpatialDerivatives, writing DP = L v^2/b (DM) to e_dp (e_dm)
//             instead of pSpace (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:269-291)
//   BORN_INJ  TimeStep then ForwardBornInjection_V(dV = e_dv, DP = e_ip, DM = e_im): the Born
//             source is added to the new field (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:305-334,
//             src/prop3DAcoVTIDenQ_DEO2_FDTD.h:420-452)
//   IMAGE     TimeStep then AdjointBornAccumulation_V(g = e_dp, DP = e_ip, DM = e_im), whose
//             pOld after the step is this step's input field (…Iso…h:518-546, …VTI…h:564-589)
//   BORN_COUPLED (Iso only) two wavefields in one pass: the background (pCur/pOld) takes
//             SAVE_DP's step, writing DP to e_dp, and the Born field (mCur/mOld) takes
//             BORN_INJ's step with dV = e_dv and that same DP
// pSpace/mSpace are left as they were by the fused epilogues.  Every expression is the one the
// separate kernel evaluates, in its order, annulus cells included.
enum Epilogue { EPI_NONE = 0, EPI_SAVE_DP = 1, EPI_BORN_INJ = 2, EPI_IMAGE = 3, EPI_BORN_COUPLED = 4 };

// wavefields a launch steps: p (Iso), p and m (VTI, TTI), background and Born field (coupled)
constexpr int fused_nf(int phys, int epi) { return phys == ISO && epi != EPI_BORN_COUPLED ? 1 : 2; }

template <int NF, int ZW, int NROW>
struct FusedLds {
    static constexpr int LZ = ZW + 7, LY = NROW + 7, YO = NROW - 7;
    float sp[2][NF][LY * LZ];     // p (and m) of plane x+3 with a 7-cell border, double-buffered
    float soy[4][NF][NROW * ZW];  // y-temporaries, ring of planes x .. x+3
    float soz[4][NF][YO * ZW];    // z-temporaries of the output rows, same ring
};

// Global accesses are global_load/store with a uniform plane base in SGPRs (saddr) and the
// 32-bit byte offset of the column inside the plane in a VGPR.  The offsets are passed
// through an empty asm once per step so that the compiler keeps their zero extension next to
// the accesses (hoisted out of the loop it would become a 64-bit VGPR address per field).  Plane
// indices outside [0, nx) are clamped for the address and their values replaced by 0;
// columns outside the grid use offset 0 and are zeroed the same way, so every address stays
// inside the field.
__device__ __forceinline__ const char *plane(const float *a, int xx, int nx, long nynz) {
    const int xc = xx < 0 ? 0 : (xx >= nx ? nx - 1 : xx);
    return reinterpret_cast<const char *>(a + (long)xc * nynz);
}
__device__ __forceinline__ float gld(const char *base, uint32_t off) {
    return *reinterpret_cast<const float *>(base + off);
}
__device__ __forceinline__ void gst(float v, float *a, int xx, long nynz, uint32_t off) {
    *reinterpret_cast<float *>(reinterpret_cast<char *>(a + (long)xx * nynz) + off) = v;
}

__device__ __forceinline__ uint32_t launder(uint32_t off) {
    asm volatile("" : "+v"(off));
    return off;
}
__device__ __forceinline__ float flip(float v, uint32_t sign) {
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) ^ sign);
}

// byte offset of column (yy, zz) inside a plane; above a free surface the column of the odd
// mirror p[-j] = -p[j] with sign = the sign bit; ok = false outside the grid (offset 0)
__device__ __forceinline__ uint32_t col_off(int yy, int zz, int ny, int nz, bool mirror, uint32_t &sign,
                                            bool &ok) {
    sign = 0;
    ok = false;
    if (yy < 0 || yy >= ny) return 0;
    if (zz >= 0 && zz < nz) {
        ok = true;
        return (uint32_t)(yy * nz + zz) * 4u;
    }
    if (mirror && zz < 0 && -zz < nz) {
        sign = 0x80000000u;
        ok = true;
        return (uint32_t)(yy * nz - zz) * 4u;
    }
    return 0;
}

__device__ __forceinline__ float sten8(const float *a) {
    // a[0..7] = values at offsets -3..+4 (D+) or -4..+3 (D-) of the centre pair a[3],a[4]
    return C1 * (-a[3] + a[4]) + C2 * (-a[2] + a[5]) + C3 * (-a[1] + a[6]) + C4 * (-a[0] + a[7]);
}

template <int N>
using IC = std::integral_constant<int, N>;

// FST: this tile holds the free-surface rows (fs and the first z-tile)
template <int PHYS, int ZW, int NROW, int EPI, bool FST>
__device__ __forceinline__ void fused_body(const FusedArgs<PHYS> &A, FusedLds<fused_nf(PHYS, EPI), ZW, NROW> &S,
                                           int zt, int yt, int xc) {
    constexpr bool fst = FST;
    constexpr int NF = fused_nf(PHYS, EPI);
    static_assert(EPI != EPI_BORN_COUPLED || PHYS == ISO, "the coupled Born step is Iso's");
    constexpr int NM = PHYS == ISO ? 1 : (PHYS == VTI ? 4 : 8);  // material fields of the sandwich
    // output operands: v, q, pOld (, mOld), then for the epilogues (COUPLED) dV, (BORN_INJ) dV,
    // DP (, DM) or (IMAGE) g, DP (, DM).  The epilogues' b is the update's: the ring bq holds b
    // at every written cell (the material is loaded there too)
    constexpr int NEX = EPI == EPI_NONE || EPI == EPI_SAVE_DP ? 0 : (EPI == EPI_BORN_COUPLED ? 1 : 1 + NF);
    constexpr int NO = 2 + NF + NEX;
    constexpr int IB = 2 + NF;  // index of the first epilogue operand
    constexpr int ZO = ZW - 7, YO = NROW - 7, LZ = ZW + 7;
    constexpr int NB = 7 * (LZ + NROW);  // border cells of the staged plane
    static_assert(NB <= ZW * NROW, "one border cell per thread at most");
    const Grid &g = A.g;
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz;
    const int t = threadIdx.x, l = t % ZW, r = t / ZW;
    const int z0 = zt * ZO, y0 = yt * YO;
    const int z = z0 - 4 + l, y = y0 - 4 + r;
    int xs, xe;
    chunk_range(A, xc, xs, xe);
    const int zlo = fst ? 0 : 4;
    const bool yzU = y >= 4 && y < ny - 4 && z >= zlo && z < nz - 4;
    // outputs: the inner (ZW - 7) x (NROW - 7) cells; the last tile along an axis also writes the
    // grid's last cells that lie in its upper halo -- annulus cells (fill_fused counts the tiles
    // so), whose new values need no D- (L = 0 there), only their own operands
    const bool out_thread = l >= 4 && (l < 4 + ZO || zt == A.nzt - 1) && r >= 4 &&
                            (r < 4 + YO || yt == A.nyt - 1) && y < ny && z < nz;
    uint32_t own_sign;
    bool own_ok;
    const uint32_t own = col_off(y, z, ny, nz, fst, own_sign, own_ok);
    // material: used inside the update region, and (the epilogues) its b at every written cell
    const uint32_t moff = (yzU || (EPI != EPI_NONE && out_thread)) ? own : 0u;
    const uint32_t ooff = out_thread ? own : 0u;
    // this thread's border cell of the staged plane, if any
    uint32_t boff = 0, bsign = 0;
    bool b_ok = false;
    int blds = 0;
    const bool has_b = t < NB;
    if (has_b) {
        int sy, sz;
        if (t < 3 * LZ) {
            sy = t / LZ;
            sz = t - sy * LZ;
        } else if (t < 3 * LZ + 7 * NROW) {
            const int m = t - 3 * LZ, c = m % 7;
            sy = 3 + m / 7;
            sz = c < 3 ? c : ZW + c;
        } else {
            const int m = t - 3 * LZ - 7 * NROW;
            sy = 3 + NROW + m / LZ;
            sz = m - (sy - 3 - NROW) * LZ;
        }
        boff = col_off(y0 - 7 + sy, z0 - 7 + sz, ny, nz, fst, bsign, b_ok);
        blds = sy * LZ + sz;
    }
    const float *fld[2] = {A.pCur, A.mCur};
    const float *mat[8] = {A.b, A.eps, A.eta, A.f, A.sT, A.cT, A.sP, A.cP};
    const float *oin[NO];
    oin[0] = A.v;
    oin[1] = A.q;
    oin[2] = A.pOld;
    if constexpr (NF == 2) oin[3] = A.mOld;
    if constexpr (EPI == EPI_BORN_COUPLED) oin[IB] = A.e_dv;
    if constexpr (EPI == EPI_BORN_INJ || EPI == EPI_IMAGE) {
        oin[IB] = EPI == EPI_BORN_INJ ? A.e_dv : A.e_dp;
        oin[IB + 1] = A.e_ip;
        if constexpr (NF == 2) oin[IB + 2] = A.e_im;
    }

    const long nynz = g.nynz;
    // this column's sponge distance in y and z (k_setup_q's rule); the x part is added per plane.
    // Iso and VTI only: TTI's step runs at the register limit, and the extra live value and
    // branch cost it more (12.7 -> 13.0 ms at 800^3) than the table saves
    constexpr bool QT = PHYS != TTI;
    const int kyz = QT ? min(min(y, ny - 1 - y), g.fs ? nz - 1 - z : min(z, nz - 1 - z)) : 0;
    // loads are issued raw; the zeroing outside the grid and the mirror's sign are applied
    // when the value is consumed, so that no wait follows a load
    auto ld = [&](const float *a, int xx, uint32_t off) -> float { return gld(plane(a, xx, nx, nynz), off); };
    // output operand k of plane xx: Q from the table while it is the profile (the index clamped
    // into the table for columns and planes whose value is never used)
    auto opnd = [&](int k, int xx, uint32_t off) -> float {
        if (QT && k == 1 && A.qtab) {
            int ks = min(min(xx, nx - 1 - xx), kyz);
            ks = ks < 0 ? 0 : (ks > A.nsponge ? A.nsponge : ks);
            return A.qtab[ks];
        }
        return ld(oin[k], xx, off);
    };
    uint32_t own_s, boff_s, moff_s, ooff_s;  // this step's copies of the offsets
    auto launder_all = [&]() {
        own_s = launder(own);
        boff_s = launder(boff);
        moff_s = launder(moff);
        ooff_s = launder(ooff);
    };
    auto fixp = [&](float v, int xx) -> float {
        v = (own_ok && xx >= 0 && xx < nx) ? v : 0.0f;
        if constexpr (FST) v = flip(v, own_sign);
        return v;
    };
    auto fixb = [&](float v, int xx) -> float {
        v = (b_ok && xx >= 0 && xx < nx) ? v : 0.0f;
        if constexpr (FST) v = flip(v, bsign);
        return v;
    };

    float pq[NF][8], pn[NF], oxq[NF][8], bq[4], bn[NF];
    float mc[NM], mnx[NM], oc[NO], onx[NO];
    const int xfirst = xs - 7;
#pragma unroll
    for (int fi = 0; fi < NF; fi++) {
#pragma unroll
        for (int j = 0; j < 8; j++) oxq[fi][j] = 0.0f;
        pq[fi][0] = 0.0f;
    }
    launder_all();
#pragma unroll
    for (int fi = 0; fi < NF; fi++) {
#pragma unroll
        for (int j = 1; j < 8; j++) pq[fi][j] = fixp(ld(fld[fi], xfirst + j - 1, own_s), xfirst + j - 1);
        pn[fi] = ld(fld[fi], xfirst + 7, own_s);
        bn[fi] = ld(fld[fi], xfirst + 3, boff_s);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) bq[j] = 0.0f;
#pragma unroll
    for (int k = 0; k < NM; k++) mnx[k] = ld(mat[k], xfirst + 3, moff_s);
#pragma unroll
    for (int k = 0; k < NO; k++) onx[k] = opnd(k, xfirst, ooff_s);

    auto iter = [&](auto U, int x) {
        constexpr int u = decltype(U)::value;  // x - xfirst = u (mod 4)
        constexpr int s2 = (u + 3) & 1;        // p-plane buffer of plane x+3
        constexpr int sw = (u + 3) & 3;        // ring slot of plane x+3 (written now)
        constexpr int sr = u & 3;              // ring slot of plane x (read now)
        // ---- 1. shift the queues, stage plane x+3 in LDS --------------------------------
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
#pragma unroll
            for (int j = 0; j < 7; j++) pq[fi][j] = pq[fi][j + 1];
            pq[fi][7] = fixp(pn[fi], x + 7);
            S.sp[s2][fi][(r + 3) * LZ + l + 3] = pq[fi][3];
            if (has_b) S.sp[s2][fi][blds] = fixb(bn[fi], x + 3);
        }
#pragma unroll
        for (int k = 0; k < NM; k++) mc[k] = mnx[k];
#pragma unroll
        for (int k = 0; k < NO; k++) oc[k] = onx[k];
        // ---- 2. request the operands of step x+1 ---------------------------------------
        launder_all();
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
            pn[fi] = ld(fld[fi], x + 8, own_s);
            bn[fi] = ld(fld[fi], x + 4, boff_s);
        }
#pragma unroll
        for (int k = 0; k < NM; k++) mnx[k] = ld(mat[k], x + 4, moff_s);
#pragma unroll
        for (int k = 0; k < NO; k++) onx[k] = opnd(k, x + 1, ooff_s);
        __syncthreads();
        // ---- 3. gradient + material sandwich at (x+3, y, z) ----------------------------
        const int xp = x + 3;
        float ox[NF], oy[NF], oz[NF];
#pragma unroll
        for (int fi = 0; fi < NF; fi++) ox[fi] = oy[fi] = oz[fi] = 0.0f;
        if (xp >= 4 && xp < nx - 4 && yzU) {
            float dx[NF], dy[NF], dz[NF];
#pragma unroll
            for (int fi = 0; fi < NF; fi++) {
                float a[8];
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = S.sp[s2][fi][(r + j) * LZ + l + 3];  // rows y-3 .. y+4
                dy[fi] = g.idy * sten8(a);
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = S.sp[s2][fi][(r + 3) * LZ + l + j];  // lanes z-3 .. z+4
                dz[fi] = g.idz * sten8(a);
                dx[fi] = g.idx * sten8(pq[fi]);
                if (fst && z == 0) {  // free-surface row: x/y derivatives vanish
                    dx[fi] = 0.0f;
                    dy[fi] = 0.0f;
                }
            }
            if constexpr (PHYS == ISO) {
#pragma unroll
                for (int fi = 0; fi < NF; fi++) {
                    ox[fi] = mc[0] * dx[fi];
                    oy[fi] = mc[0] * dy[fi];
                    oz[fi] = mc[0] * dz[fi];
                }
            } else if constexpr (PHYS == VTI) {
                vti_material(dx[0], dy[0], dz[0], dx[1], dy[1], dz[1], mc[1], mc[2], mc[3], mc[0], ox[0], oy[0],
                             oz[0], ox[1], oy[1], oz[1]);
            } else {
                tti_material(dx[0], dy[0], dz[0], dx[1], dy[1], dz[1], mc[4], mc[5], mc[6], mc[7], mc[1], mc[2],
                             mc[3], mc[0], ox[0], oy[0], oz[0], ox[1], oy[1], oz[1]);
            }
        }
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
#pragma unroll
            for (int j = 0; j < 7; j++) oxq[fi][j] = oxq[fi][j + 1];
            oxq[fi][7] = ox[fi];
            S.soy[sw][fi][r * ZW + l] = oy[fi];
            if (r >= 4 && r < 4 + YO) S.soz[sw][fi][(r - 4) * ZW + l] = oz[fi];
        }
        bq[sw] = mc[0];
        // ---- 4. D- and leapfrog at plane x (temporaries of x written three steps ago) ----
        if (x < xs || x >= xe || !out_thread) return;
        const uint32_t so = launder(ooff);  // zero extension next to the stores
        const bool upd = x >= 4 && x < nx - 4 && yzU;
        float L[NF];
#pragma unroll
        for (int fi = 0; fi < NF; fi++) L[fi] = 0.0f;
        if (upd) {
#pragma unroll
            for (int fi = 0; fi < NF; fi++) {
                if (FST && z == 0) {
                    L[fi] = 0.0f + 0.0f + 0.0f;
                    continue;
                }
                float a[8];
                const float dxm = g.idx * sten8(oxq[fi]);
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = S.soy[sr][fi][(r - 4 + j) * ZW + l];  // rows y-4 .. y+3
                const float dym = g.idy * sten8(a);
                if (FST && z < 4) {
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const int zz = z - 4 + j;  // even mirror: t[-j] = t[j-1]; lane of kz is kz+4
                        const int zm = zz >= 0 ? zz : -zz - 1;
                        a[j] = S.soz[sr][fi][(r - 4) * ZW + zm + 4];
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 8; j++) a[j] = S.soz[sr][fi][(r - 4) * ZW + l - 4 + j];  // lanes z-4 .. z+3
                }
                const float dzm = g.idz * sten8(a);
                L[fi] = dxm + dym + dzm;
            }
        }
        const float V = oc[0];
        // new field values (pOld's buffer): updated inside the region, kept in the annulus
        float pn_[NF];
        if (upd) {
            const float dt2V2_B = A.dt2 * V * V / bq[sr];
            const float qk = oc[1];
            pn_[0] = leapfrog(dt2V2_B, L[0], qk, pq[0][0], oc[2]);
            if constexpr (NF == 2) pn_[1] = leapfrog(dt2V2_B, L[1], qk, pq[1][0], oc[3]);
        } else {
#pragma unroll
            for (int fi = 0; fi < NF; fi++) pn_[fi] = oc[2 + fi];
        }
        if constexpr (EPI == EPI_NONE) {
            // annulus: the spatial output is zeroed, the wavefield is not updated
            if (upd) {
                gst(pn_[0], A.pOld, x, nynz, so);
                if constexpr (NF == 2) gst(pn_[1], A.mOld, x, nynz, so);
            }
            if (A.pSpace) {
                gst(L[0], A.pSpace, x, nynz, so);
                if constexpr (NF == 2) gst(L[1], A.mSpace, x, nynz, so);
            }
        } else if constexpr (EPI == EPI_SAVE_DP) {
            if (upd) {
                gst(pn_[0], A.pOld, x, nynz, so);
                if constexpr (NF == 2) gst(pn_[1], A.mOld, x, nynz, so);
            }
            const float sc = V * V / bq[sr];  // scaleSpatialDerivatives: pSpace *= v^2/b
            gst(L[0] * sc, A.e_dp, x, nynz, so);
            if constexpr (NF == 2) gst(L[1] * sc, A.e_dm, x, nynz, so);
        } else if constexpr (EPI == EPI_BORN_INJ) {
            const float B = bq[sr], dv = oc[IB];
            if constexpr (PHYS == ISO) {
                const float dt2v2OverB = A.dt * A.dt * V * V / B;
                const float factorV = 2 * B * dv / (V * V * V);
                pn_[0] += dt2v2OverB * oc[IB + 1] * factorV;
            } else {
                const float factor = 2 * A.dt * A.dt * dv / V;
                pn_[0] += factor * oc[IB + 1];
                pn_[1] += factor * oc[IB + 2];
            }
            gst(pn_[0], A.pOld, x, nynz, so);
            if constexpr (NF == 2) gst(pn_[1], A.mOld, x, nynz, so);
        } else if constexpr (EPI == EPI_BORN_COUPLED) {
            // background: SAVE_DP's step; Born field: BORN_INJ's step with the DP just formed
            if (upd) gst(pn_[0], A.pOld, x, nynz, so);
            const float B = bq[sr], dv = oc[IB];
            const float sc = V * V / B;
            const float DP = L[0] * sc;
            gst(DP, A.e_dp, x, nynz, so);
            const float dt2v2OverB = A.dt * A.dt * V * V / B;
            const float factorV = 2 * B * dv / (V * V * V);
            pn_[1] += dt2v2OverB * DP * factorV;
            gst(pn_[1], A.mOld, x, nynz, so);
        } else {  // EPI_IMAGE: g += 2b/v^3 (DP p + DM m) with this step's input fields
            if (upd) {
                gst(pn_[0], A.pOld, x, nynz, so);
                if constexpr (NF == 2) gst(pn_[1], A.mOld, x, nynz, so);
            }
            const float B = bq[sr];
            float gk = oc[IB];
            if constexpr (PHYS == ISO) {
                const float factorV = +2 * B / (V * V * V);
                gk += factorV * oc[IB + 1] * pq[0][0];
            } else {
                const float factor = 2 * B / (V * V * V);
                gk += factor * (oc[IB + 1] * pq[0][0] + oc[IB + 2] * pq[1][0]);
            }
            gst(gk, A.e_dp, x, nynz, so);
        }
    };

    // steps x < xs only fill the queues; steps x >= xe (at most 3, to keep the 4-step
    // unrolling) store nothing.  Values these steps load from planes that neighbouring chunks
    // update (pOld, mOld) are never used.
    for (int xb = xfirst; xb < xe; xb += 4) {
        iter(IC<0>{}, xb);
        iter(IC<1>{}, xb + 1);
        iter(IC<2>{}, xb + 2);
        iter(IC<3>{}, xb + 3);
    }
}

// tile index (in the XCD-local order) -> (zt, yt, xc)
template <int PHYS>
__device__ __forceinline__ void tile_coords(const FusedArgs<PHYS> &A, int tile, int &zt, int &yt, int &xc) {
    const int per_xc = A.nzt * A.nyt;
    xc = tile / per_xc;
    const int rem = tile - xc * per_xc;
    if (A.bz == 0) {  // z fastest, then y
        zt = rem % A.nzt;
        yt = rem / A.nzt;
    } else if (A.bz < 0) {  // y fastest, then z
        yt = rem % A.nyt;
        zt = rem / A.nyt;
    } else {
        // rows of by y-tiles; inside a row, blocks of bz x (rows in the row) tiles, the last
        // row and the last block of a row ragged
        const int yb = rem / (A.by * A.nzt), r2 = rem - yb * A.by * A.nzt;
        const int hy = min(A.by, A.nyt - yb * A.by);
        const int zb = r2 / (A.bz * hy), r3 = r2 - zb * A.bz * hy;
        const int wz = min(A.bz, A.nzt - zb * A.bz);
        zt = zb * A.bz + r3 % wz;
        yt = yb * A.by + r3 / wz;
    }
}

// x-planes [xs, xe) of chunk xc
template <int PHYS>
__device__ __forceinline__ void chunk_range(const FusedArgs<PHYS> &A, int xc, int &xs, int &xe) {
    xs = xc * A.cx + A.cxu * min(xc, A.nlong);
    xe = xc == A.nxc - 1 ? (int)A.g.nx : xs + A.cx + (xc < A.nlong ? A.cxu : 0);
}

// workgroup bid of a launch of nwg workgroups over A's tiles -> (zt, yt, xc); false for the
// padding workgroups of a batch (nwg rounded up to a multiple of 8).  XCD-aware numbering:
// workgroup id bid runs on XCD bid mod 8; each XCD owns a contiguous run of the tile order, so
// that the workgroups resident on one XCD at a time are neighbours and share their halo lines
// in its L2.
template <int PHYS>
__device__ __forceinline__ bool fused_tile_of_wg(const FusedArgs<PHYS> &A, int bid, int &loc, int &zt, int &yt,
                                                 int &xc) {
    const int xcd = bid & 7;
    const int ntiles = A.nzt * A.nyt * A.nxc;
    if (A.xmap > 0) {
        // groups of G consecutive tiles dealt to the XCDs in turn: the eight XCDs' rounds run
        // neighbouring groups at the same time; ids past the last tile do nothing
        const int G = A.xmap;
        loc = bid >> 3;
        const int sr = loc / G;
        const int tile = (sr * 8 + xcd) * G + (loc - sr * G);
        if (tile >= ntiles) return false;
        tile_coords(A, tile, zt, yt, xc);
        return true;
    }
    const int q8 = ntiles >> 3, r8 = ntiles & 7;
    const int start = xcd < r8 ? xcd * (q8 + 1) : r8 * (q8 + 1) + (xcd - r8) * q8;
    const int count = q8 + (xcd < r8 ? 1 : 0);
    loc = bid >> 3;  // position in this XCD's run of tiles
    if (loc >= count) return false;
    tile_coords(A, start + loc, zt, yt, xc);
    return true;
}

// zeroes the per-XCD round counters before a launch with rounds in step
__global__ void k_zero_u32(unsigned *a, int n) {
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0u;
}

template <int PHYS, int ZW, int NROW, int EPI>
__global__ void __launch_bounds__(ZW *NROW) k_fused_step(FusedArgs<PHYS> A) {
    __shared__ FusedLds<fused_nf(PHYS, EPI), ZW, NROW> S;
    const int bid = (int)blockIdx.x, xcd = bid & 7;
    int loc, zt, yt, xc;
    if (!fused_tile_of_wg(A, bid, loc, zt, yt, xc)) return;
    const int per = A.per_xcd;
    if (per > 0 && loc >= per) {
        // Rounds in step: workgroups start on an XCD in id order, one per CU, so this
        // workgroup belongs to round loc / per.  It waits until the XCD has finished every
        // earlier round, so that a round's neighbouring tiles march through x together and
        // read their shared halo lines while these are in the XCD's L2.  Those workgroups
        // were started before this one; the wait is bounded and orders nothing the results
        // depend on.
        if (threadIdx.x == 0) {
            const unsigned target = (unsigned)(loc / per * per);
            for (int i = 0; i < 4096; i++) {
                if (__hip_atomic_load(&A.sync[xcd], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
                __builtin_amdgcn_s_sleep(16);
            }
        }
        __syncthreads();
    }
    if (A.g.fs && zt == 0)
        fused_body<PHYS, ZW, NROW, EPI, true>(A, S, zt, yt, xc);
    else
        fused_body<PHYS, ZW, NROW, EPI, false>(A, S, zt, yt, xc);
    if (per > 0) {
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_fetch_add(&A.sync[xcd], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Several grids (shots) stepped by one launch: grid s owns workgroups [wg[s], wg[s+1]), each
// range a multiple of 8 long so that the XCD-aware numbering holds inside every range.  The
// grids share the physics, the tile shape and the epilogue; sizes and arrays are their own.
constexpr int FUSED_MAX_BATCH = 8;
template <int PHYS>
struct FusedBatch {
    FusedArgs<PHYS> a[FUSED_MAX_BATCH];
    int wg[FUSED_MAX_BATCH + 1];
    int n;
};

template <int PHYS, int ZW, int NROW, int EPI>
__global__ void __launch_bounds__(ZW *NROW) k_fused_batch(FusedBatch<PHYS> B) {
    __shared__ FusedLds<fused_nf(PHYS, EPI), ZW, NROW> S;
    const int bid = (int)blockIdx.x;
    int s = 0;
    while (s + 1 < B.n && bid >= B.wg[s + 1]) s++;
    const FusedArgs<PHYS> &A = B.a[s];
    int loc, zt, yt, xc;
    if (!fused_tile_of_wg(A, bid - B.wg[s], loc, zt, yt, xc)) return;
    if (A.g.fs && zt == 0)
        fused_body<PHYS, ZW, NROW, EPI, true>(A, S, zt, yt, xc);
    else
        fused_body<PHYS, ZW, NROW, EPI, false>(A, S, zt, yt, xc);
}

// x-planes per chunk: balance whole rounds of workgroups (per_cu per CU) against the 7-step
// warm-up and the padding to a multiple of `unroll` steps that each chunk pays
// nx planes in nxc chunks for a kernel whose x-loop runs in steps of u (7 warm-up steps, then one
// step per plane, padded to a multiple of u): lengths of the form k u + 1 need no padding, so every
// chunk gets the same such length b, nlong of them u more, and the last one the remainder.
// 800 planes in 8 chunks of an 8-step loop: 5 x 97 + 3 x 105 = 856 steps, against 8 x 112 = 896
// for 8 x 100
static inline void chunk_split(long nx, long nxc, int u, int &b, int &nlong) {
    const long per = nx / nxc;
    const long bb = per >= 1 ? (per - 1) / u * u + 1 : 1;
    b = (int)bb;
    nlong = (int)std::min(nxc - 1, (nx - nxc * bb) / u);
}

// steps of the x-loop of the longest chunk of that split
static inline long chunk_steps(long nx, long nxc, int u) {
    int b, nlong;
    chunk_split(nx, nxc, u, b, nlong);
    const long last = nx - (nxc - 1) * (long)b - (long)u * nlong;
    const long longest = std::max<long>(nlong > 0 ? b + u : b, last);
    return (longest + 7 + u - 1) / u * u;
}

static inline int fused_cx(const Prop *p, long tiles, int unroll = 4, int per_cu = 1) {
    if (p->fused_chunk > 0) return (int)std::min<long>(p->fused_chunk, p->nx);
    if (const char *e = std::getenv("WAVEFD_FUSED_CX")) {
        const long v = std::atol(e);
        if (v > 0) return (int)std::min<long>(v, p->nx);
    }
    const long cus = (p->cu_count > 0 ? p->cu_count : 256) * per_cu;
    long best_cx = p->nx;
    double best = 1e300;
    for (long nxc = 1; nxc <= 64 && nxc <= p->nx; nxc++) {
        const long cx = (p->nx + nxc - 1) / nxc;
        if ((p->nx + cx - 1) / cx != nxc) continue;  // no length gives nxc chunks
        const long steps = chunk_steps(p->nx, nxc, unroll);
        const long rounds = (tiles * nxc + cus - 1) / cus;
        const double cost = (double)rounds * (double)steps;
        if (cost < best * 0.999) {
            best = cost;
            best_cx = cx;
        }
    }
    return (int)best_cx;
}

// tiles along z (y) for o outputs per tile: enough that the outputs cover the update region
// [.., n - 4) and the last tile's upper halo (3 cells) reaches the grid's end, i.e. o * tiles >=
// n - 3: the cells past the outputs are then annulus cells, which the last tile writes from its
// halo lanes (rows).  Against ceil(n / o): 800 cells take 14 tiles of 57 (not 15), 47 of 17 (not 48)
static inline long fused_tiles_along(long n, int o) { return std::max(1L, (n - 3 + o - 1) / o); }

// e0, e1, e2: the epilogue's arrays (device), see Epilogue: SAVE_DP (DP, DM, -),
// BORN_INJ (dV, DP, DM), IMAGE (g, DP, DM).  Returns the number of tiles.
template <int PHYS, int ZW, int NROW, int EPI>
static inline long fill_fused(Prop *p, bool linear, float *e0, float *e1, float *e2, FusedArgs<PHYS> &A,
                              int unroll = 4, int (*cx_rule)(const Prop *, long, int, int) = nullptr) {
    constexpr int ZO = ZW - 7, YO = NROW - 7;
    const Field *f = p->f;
    A.g = p->g;
    A.dt2 = p->dt * p->dt;
    A.dt = p->dt;
    A.e_dp = A.e_dm = nullptr;
    A.e_dv = A.e_ip = A.e_im = nullptr;
    if (EPI == EPI_SAVE_DP) {
        A.e_dp = e0;
        A.e_dm = e1;
    } else if (EPI == EPI_BORN_INJ) {
        A.e_dv = e0;
        A.e_ip = e1;
        A.e_im = e2;
    } else if (EPI == EPI_IMAGE) {
        A.e_dp = e0;
        A.e_ip = e1;
        A.e_im = e2;
    } else if (EPI == EPI_BORN_COUPLED) {
        A.e_dv = e0;
        A.e_dp = e1;
    }
    A.nzt = (int)fused_tiles_along(p->nz, ZO);
    A.nyt = (int)fused_tiles_along(p->ny, YO);
    A.cx = cx_rule ? cx_rule(p, (long)A.nzt * A.nyt, PHYS, EPI) : fused_cx(p, (long)A.nzt * A.nyt, unroll);
    // order of the tiles in an XCD's run: VTI in blocks of (CUs per XCD) tiles, 4 along z; Iso
    // and TTI plain z-fastest, so that the tiles resident on an XCD hold whole z-columns
    // (measured on MI355X, plain against 4 x 8 blocks: TTI 800^3 10.67 against 11.55 ms, 512^3
    // 3.05 against 3.23 ms, Iso 512^3 1.135 against 1.147 ms; VTI 800^3 8.97 against 8.79 ms).
    // WAVEFD_FUSED_BZ=k: blocks of k along z; WAVEFD_FUSED_ORDER=0: plain, 2: y-fastest (for measurements)
    const int xcd_cus = std::max(1, p->cu_count / 8);
    int bz = PHYS == VTI ? 4 : 0;
    if (const char *e = std::getenv("WAVEFD_FUSED_BZ")) bz = std::max(0, std::atoi(e));
    if (const char *e = std::getenv("WAVEFD_FUSED_ORDER")) {
        if (std::atoi(e) == 0) bz = 0;
        if (std::atoi(e) == 2) bz = -1;
    }
    A.bz = bz > 0 ? std::min(bz, A.nzt) : bz;
    A.by = A.bz > 0 ? std::max(1, xcd_cus / A.bz) : 1;
    const long nxc = (p->nx + A.cx - 1) / A.cx;
    A.nxc = (int)nxc;
    A.cxu = unroll;
    chunk_split(p->nx, nxc, unroll, A.cx, A.nlong);
    A.pCur = f[F_PCUR].d;
    A.pOld = f[F_POLD].d;
    A.pSpace = linear ? nullptr : f[F_PSPACE].d;
    A.mCur = f[F_MCUR].d;
    A.mOld = f[F_MOLD].d;
    A.mSpace = linear ? nullptr : f[F_MSPACE].d;
    A.v = f[F_V].d;
    A.b = f[F_B].d;
    A.q = f[F_Q].d;
    A.qtab = (PHYS != TTI && p->q_profile) ? p->qtab : nullptr;
    A.nsponge = (int)p->nsponge;
    A.eps = f[F_EPS].d;
    A.eta = f[F_ETA].d;
    A.f = f[F_F].d;
    A.sT = f[F_ST].d;
    A.cT = f[F_CT].d;
    A.sP = f[F_SP].d;
    A.cP = f[F_CP].d;
    A.per_xcd = 0;
    A.xmap = 0;
    A.slack = -1;  // the launch's default
    if (const char *e = std::getenv("WAVEFD_FUSED_SLACK")) A.slack = std::max(0, std::atoi(e));
    A.sync = p->fused_sync;
    return (long)A.nzt * A.nyt * nxc;
}

// rounds in step: the automatic rule, overridden by the handle's setting, else by the
// environment; never where the rounds cannot be formed (min_rounds of one workgroup per CU: 20
// for k_fused_step, 12 for the wide tile, whose XCDs take turns)
static inline bool fused_lockstep_of(const Prop *p, int phys, int epi, long ntiles, long min_rounds = 20) {
    const bool can = p->cu_count % 8 == 0 && ntiles >= min_rounds * p->cu_count;
    bool lockstep = can && phys != ISO && epi == EPI_NONE;
    if (p->fused_lockstep >= 0)
        lockstep = can && p->fused_lockstep != 0;
    else if (const char *e = std::getenv("WAVEFD_FUSED_LOCKSTEP"))
        lockstep = can && std::atoi(e) != 0;
    return lockstep;
}

template <int PHYS, int ZW, int NROW, int EPI>
static inline bool launch_fused(Prop *p, bool linear, float *e0 = nullptr, float *e1 = nullptr,
                                float *e2 = nullptr) {
    FusedArgs<PHYS> A;
    const long ntiles = fill_fused<PHYS, ZW, NROW, EPI>(p, linear, e0, e1, e2, A);
    // rounds in step when there are several rounds of one workgroup per CU (the kernel's LDS
    // and registers allow no second one per CU).  It pays where the halo re-reads are heavy
    // (VTI, TTI: 4 and 8 material fields; measured -20 % fetched bytes, -3 % time for TTI
    // 800^3); for Iso, and for the Born epilogues (VTI Born passes on 200^3..600^3 windows:
    // 6876 against 7032 shots per hour), the waiting costs more than the shared lines save.  It
    // needs many rounds: 800^3 (21 rounds of 256 workgroups) TTI 13.31 -> 12.73 ms, VTI 10.36 ->
    // 10.14 ms; 512^3 (18 rounds) TTI 3.43 -> 3.70 ms, VTI 2.80 -> 2.98 ms.
    const bool lockstep = fused_lockstep_of(p, PHYS, EPI, ntiles);
    if (lockstep && p->fused_sync) {
        A.per_xcd = p->cu_count / 8;
        if (A.slack < 0) A.slack = 0;
        // a kernel, not a memset, so that the step can be captured in a graph
        k_zero_u32<<<1, 64, 0, p->stream>>>(p->fused_sync, 8);
        WFD_LAUNCHED();
    }
    k_fused_step<PHYS, ZW, NROW, EPI><<<dim3((unsigned)ntiles), dim3(ZW * NROW), 0, p->stream>>>(A);
    return true;
}

}  // namespace wfd

#include "wfd_wide.hpp"

namespace wfd {

static inline bool fused_fits(Prop *p) {
    // plane offsets are 32-bit byte offsets
    if (p->ny * p->nz >= (1L << 29)) {
        set_error("fused step: a y-z plane must hold fewer than 2^29 cells");
        return false;
    }
    return true;
}

// tile shapes of the fused kernel (temporaries along z x along y; outputs (z-7) x (y-7)):
// 64 x 16 and 32 x 32 with one thread per cell (k_fused_step), 64 x 32 with four y-rows per
// thread for Iso, 64 x 24 with three for VTI and TTI (k_wide_step); 64 x 16 with two y-rows per
// thread and two workgroups per CU for the one-wavefield steps (Iso's; the others take tile 3 for
// it)
enum FusedTile { TILE_AUTO = 0, TILE_64x16 = 1, TILE_32x32 = 2, TILE_64x32W = 3, TILE_64x16W = 4 };

static inline bool set_fused_tile(Prop *p, long tile) {
    if (tile < TILE_AUTO || tile > TILE_64x16W) {
        set_error("SetFusedTile: tile must be 0 (automatic), 1 (64x16), 2 (32x32), 3 (wide: 64x32 Iso, 64x24 "
                  "VTI/TTI) or 4 (wide 64x16, two workgroups per CU: Iso's one-wavefield steps)");
        return false;
    }
    p->fused_tile = (int)tile;
    return true;
}

// automatic choice (measured on MI355X, 800^3 TTI, 512^3 VTI, 256^3 and 512^3 Iso).  VTI and TTI
// steps: the wide tile (3 y-rows per thread, 57 x 17 outputs), whose straight-line rows give each
// wave the instruction-level parallelism the 64 x 16 tile gets from more waves (VTI 512^3 2.34
// against 2.72 ms, TTI 800^3 12.39 against 12.69 ms; the Born epilogues, 24 VTI shots on
// 200^3..600^3 windows: 7938 shots per hour against 7300 (64 x 16) and 6768 (32 x 32)).  Iso's
// one-wavefield steps: the two-row wide tile, two workgroups per CU (512^3 1.097 against 1.328 ms
// for the 32 x 32 tile and 1.50 for 64 x 32 wide, 256^3 0.174 against 0.179 ms; the 24 Iso Born
// shots 8299 against 8000 shots per hour with the wide tile's epilogues, the separate calls 6184
// against 5583).  The coupled Iso Born step (two wavefields, 167 VGPRs at two rows: one workgroup
// per CU) keeps the wide tile.
static inline int fused_tile_of(const Prop *p, int epi) {
    if (p->fused_tile == TILE_64x16W && fused_nf(p->phys, epi) != 1) return TILE_64x32W;
    if (p->fused_tile != TILE_AUTO) return p->fused_tile;
    return fused_nf(p->phys, epi) == 1 ? TILE_64x16W : TILE_64x32W;
}

// the fused step's launch configuration (tile, rounds in step, x-chunk length); results do not
// depend on it.  tile 0..4 as SetFusedTile, lockstep -1 automatic / 0 / 1, chunk 0 automatic.
static inline bool set_fused_config(Prop *p, long tile, long lockstep, long chunk) {
    if (lockstep < -1 || lockstep > 1 || chunk < 0) {
        set_error("SetFusedConfig: lockstep must be -1, 0 or 1 and chunk >= 0");
        return false;
    }
    if (!set_fused_tile(p, tile)) return false;
    p->fused_lockstep = (int)lockstep;
    p->fused_chunk = chunk;
    return true;
}

// the configuration a plain TimeStep launches with: out[0] tile (1 to 4), out[1] lockstep
// (0 or 1), out[2] x-planes per chunk, out[3] workgroups
static inline bool get_fused_config(const Prop *p, long *out) {
    if (!out) {
        set_error("GetFusedConfig: null output");
        return false;
    }
    const int tile = fused_tile_of(p, EPI_NONE);
    const int zo = (tile == TILE_32x32 ? 32 : 64) - 7;
    const int yo = (tile == TILE_32x32 ? 32 : (tile == TILE_64x32W ? wide_rows(p->phys, EPI_NONE) : 16)) - 7;
    const long tiles = fused_tiles_along(p->nz, zo) * fused_tiles_along(p->ny, yo);
    const long cx = tile == TILE_64x32W   ? wide_cx(p, tiles, p->phys, EPI_NONE)
                    : tile == TILE_64x16W ? wide_cx2(p, tiles, p->phys, EPI_NONE)
                                          : fused_cx(p, tiles, 4);
    const long ntiles = tiles * ((p->nx + cx - 1) / cx);
    out[0] = tile;
    out[1] = fused_lockstep_of(p, p->phys, EPI_NONE, ntiles, tile == TILE_64x32W ? WIDE_LOCKSTEP_ROUNDS : 20) ? 1 : 0;
    out[2] = cx;
    out[3] = ntiles;
    return true;
}

template <int PHYS, int EPI>
static inline bool launch_fused_tile(Prop *p, bool linear, float *e0, float *e1, float *e2) {
    const int tile = fused_tile_of(p, EPI);
    if constexpr (fused_nf(PHYS, EPI) == 1)
        if (tile == TILE_64x16W) return launch_wide<PHYS, EPI, 2>(p, linear, e0, e1, e2);
    if (tile == TILE_64x32W) return launch_wide<PHYS, EPI>(p, linear, e0, e1, e2);
    if (tile == TILE_32x32) return launch_fused<PHYS, 32, 32, EPI>(p, linear, e0, e1, e2);
    return launch_fused<PHYS, 64, 16, EPI>(p, linear, e0, e1, e2);
}

template <int EPI>
static inline bool step_fused_epi(Prop *p, bool linear, float *e0, float *e1, float *e2) {
    if (!fused_fits(p)) return false;
    bool ok;
    if constexpr (EPI == EPI_BORN_COUPLED) {
        if (p->phys != ISO) {
            set_error("the coupled Born step is defined for Iso only");
            return false;
        }
        ok = launch_fused_tile<ISO, EPI>(p, linear, e0, e1, e2);
    } else if (p->phys == ISO) {
        ok = launch_fused_tile<ISO, EPI>(p, linear, e0, e1, e2);
    } else if (p->phys == VTI) {
        ok = launch_fused_tile<VTI, EPI>(p, linear, e0, e1, e2);
    } else {
        ok = launch_fused_tile<TTI, EPI>(p, linear, e0, e1, e2);
    }
    if (!ok) return false;
    WFD_LAUNCHED();
    return true;
}

static inline bool step_fused(Prop *p, bool linear) {
    return step_fused_epi<EPI_NONE>(p, linear, nullptr, nullptr, nullptr);
}

// one launch of k_fused_batch over the n grids, on ps[0]'s stream (the caller orders the
// other handles' streams around it); e0/e1/e2: per-grid epilogue arrays (device) or null
template <int PHYS, int ZW, int NROW, int EPI>
static inline bool launch_fused_batch(Prop **ps, int n, float *const *e0, float *const *e1, float *const *e2) {
    FusedBatch<PHYS> B = {};
    B.n = n;
    long wg = 0;
    for (int i = 0; i < n; i++) {
        B.wg[i] = (int)wg;
        const long t = fill_fused<PHYS, ZW, NROW, EPI>(ps[i], false, e0 ? e0[i] : nullptr, e1 ? e1[i] : nullptr,
                                                      e2 ? e2[i] : nullptr, B.a[i]);
        wg += (t + 7) / 8 * 8;
        if (wg >= (1L << 31)) {
            set_error("batched step: too many workgroups");
            return false;
        }
    }
    B.wg[n] = (int)wg;
    k_fused_batch<PHYS, ZW, NROW, EPI><<<dim3((unsigned)wg), dim3(ZW * NROW), 0, ps[0]->stream>>>(B);
    return true;
}

template <int PHYS, int EPI>
static inline bool launch_fused_batch_tile(Prop **ps, int n, float *const *e0, float *const *e1, float *const *e2) {
    // the first handle's tile for all: the wide tiles (the automatic choice) through k_wide_batch,
    // the one-thread-per-cell tiles through k_fused_batch
    const int tile = fused_tile_of(ps[0], EPI);
    if constexpr (fused_nf(PHYS, EPI) == 1)
        if (tile == TILE_64x16W) return launch_wide_batch<PHYS, EPI, 2>(ps, n, e0, e1, e2);
    if (tile == TILE_64x32W) return launch_wide_batch<PHYS, EPI>(ps, n, e0, e1, e2);
    if (tile == TILE_32x32) return launch_fused_batch<PHYS, 32, 32, EPI>(ps, n, e0, e1, e2);
    return launch_fused_batch<PHYS, 64, 16, EPI>(ps, n, e0, e1, e2);
}

template <int EPI>
static inline bool step_fused_batch(Prop **ps, int n, float *const *e0, float *const *e1, float *const *e2) {
    for (int i = 0; i < n; i++)
        if (!fused_fits(ps[i])) return false;
    bool ok;
    if constexpr (EPI == EPI_BORN_COUPLED) {
        if (ps[0]->phys != ISO) {
            set_error("the coupled Born step is defined for Iso only");
            return false;
        }
        ok = launch_fused_batch_tile<ISO, EPI>(ps, n, e0, e1, e2);
    } else if (ps[0]->phys == ISO)
        ok = launch_fused_batch_tile<ISO, EPI>(ps, n, e0, e1, e2);
    else if (ps[0]->phys == VTI)
        ok = launch_fused_batch_tile<VTI, EPI>(ps, n, e0, e1, e2);
    else
        ok = launch_fused_batch_tile<TTI, EPI>(ps, n, e0, e1, e2);
    if (!ok) return false;
    WFD_LAUNCHED();
    return true;
}

}  // namespace wfd
