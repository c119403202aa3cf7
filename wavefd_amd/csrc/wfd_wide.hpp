// wfd_wide.hpp -- the fused single-pass time step on a wide tile: register-blocked in y.
//
// Same decomposition and semantics as k_fused_step (wfd_fused.hpp): 2.5D, marching along x,
// the D+ sandwich of plane x+3 and D- + leapfrog of plane x in one pass, temporaries never in
// HBM, bit-identical results (every expression in the reference's order).  What changes is
// the work per thread:
//   * a workgroup is 512 threads, 64 lanes along z by 8 waves along y, and each thread owns
//     R = 4 adjacent y-rows of its lane, so the tile is 64 x 32 temporaries with 57 x 25 outputs
//     (k_fused_step: 64 x 16 with 57 x 9).  Halo cells -- where the sandwich, and for TTI all 8
//     material fields, are recomputed -- fall from 1.00 to 0.44 per output, and the staged p
//     (m) plane with its 7-cell border from 3.2 to 1.9 cells per output;
//   * the D+y operand of a thread's R rows is one window of R + 7 staged rows, R of them the
//     thread's own values (registers), and the D-y operand one window of R + 7 ring rows;
//   * the x-loop is unrolled by 8, the period of the two 8-deep x-queues (p(x..x+7) for D+x, the
//     x-temporaries of x-4..x+3 for D-x), so a queue entry is addressed by a compile-time slot
//     and no register moves shift them;
//   * operands are requested where their registers become free: the material of plane x+4 right
//     after the sandwich of plane x+3 consumed it, the output operands of plane x+1 right after
//     the update of plane x; p(x+8) before the barrier, one step ahead.
// Waves are numbered by a scalar (readfirstlane), so every per-row condition is uniform.
//
// Resources (TTI): 8 waves, up to 256 VGPRs each, 161,040 B of LDS: one workgroup per CU.
//
// Included by wfd_fused.hpp after its tile helpers (FusedArgs, col_off, fused_tile_of_wg, ...).
#pragma once

namespace wfd {

// LDS (address space 3) pointers are 32-bit byte addresses inside the workgroup's allocation
using lds_float = __attribute__((address_space(3))) float;
__device__ __forceinline__ uint32_t lds_addr(float *generic) {
    return (uint32_t)(uintptr_t)(lds_float *)generic;
}
__device__ __forceinline__ lds_float *lds_at(uint32_t a) { return (lds_float *)(uintptr_t)a; }

template <int NF, int R>
struct WideLds {
    static constexpr int ZW = 64, NROW = 8 * R, LZ = ZW + 7, LY = NROW + 7, YO = NROW - 7;
    float sp[2][NF][LY * LZ];     // p (and m) of plane x+3 with a 7-cell border, double-buffered
    // y-temporaries, ring of planes x .. x+3; z-temporaries of the output rows, same ring.  Both
    // padded by R rows on either side, which the D- windows of a wave's non-output rows reach
    float soy[4][NF][(NROW + 2 * R) * ZW];
    float soz[4][NF][(YO + 2 * R) * ZW];
};

constexpr int WIDE_THREADS = 512;  // 64 lanes x 8 waves

// a scheduling fence: the compiler's scheduler otherwise interleaves the rows and hoists the
// operand requests of later phases, which keeps several rows' temporaries live at once and
// exceeds the 256 VGPRs of a 512-thread workgroup
#define WIDE_FENCE() __builtin_amdgcn_sched_barrier(0)

#define WIDE_FENCE_OFF()


// y-rows per thread: 4 with one wavefield; 3 with two (VTI, TTI, the coupled Iso Born step), whose
// 8-deep x-queues (and 4 or 8 material fields per row) would not fit 256 VGPRs, nor the rings
// 160 KiB of LDS, at 4 rows
constexpr int wide_r(int phys, int epi) { return fused_nf(phys, epi) == 1 ? 4 : 3; }
constexpr int wide_rows(int phys, int epi) { return 8 * wide_r(phys, epi); }
// waves per SIMD the kernel is built for: 2 (one 512-thread workgroup per CU, up to 256 VGPRs), or
// 4 for the one-wavefield tile of 2 rows per thread (two workgroups per CU, 128 VGPRs; 105 used,
// 46,856 B of LDS each)
constexpr int wide_wpe(int nf, int r) { return nf == 1 && r == 2 ? 4 : 2; }

template <int PHYS, int R, int EPI, bool FST>
__device__ __forceinline__ void wide_body(const FusedArgs<PHYS> &A, WideLds<fused_nf(PHYS, EPI), R> &S, int zt,
                                          int yt, int xc) {
    constexpr int NF = fused_nf(PHYS, EPI);
    static_assert(EPI != EPI_BORN_COUPLED || PHYS == ISO, "the coupled Born step is Iso's");
    constexpr int NM = PHYS == ISO ? 1 : (PHYS == VTI ? 4 : 8);
    // output operands as k_fused_step's: v, q, pOld (, mOld), then the epilogue's own; their b is
    // the update's (bq holds b at every written cell)
    constexpr int NEX = EPI == EPI_NONE || EPI == EPI_SAVE_DP ? 0 : (EPI == EPI_BORN_COUPLED ? 1 : 1 + NF);
    constexpr int NO = 2 + NF + NEX;
    constexpr int IB = 2 + NF;
    constexpr int ZW = 64, NROW = 8 * R, ZO = ZW - 7, YO = NROW - 7, LZ = ZW + 7;
    constexpr int NB = 7 * (LZ + NROW);  // border cells of the staged plane
    constexpr int NBT = (NB + WIDE_THREADS - 1) / WIDE_THREADS;
    const Grid &g = A.g;
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz;
    const int t = threadIdx.x, l = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);  // wave: uniform
    const int rb = R * w;                                  // tile row of this thread's first row
    const int z0 = zt * ZO, y0 = yt * YO;
    const int z = z0 - 4 + l;
    const int yb = y0 - 4 + rb;
    int xs, xe;
    chunk_range(A, xc, xs, xe);
    const int zlo = FST ? 0 : 4;
    const bool zU = z >= zlo && z < nz - 4;
    const bool zOut = l >= 4 && l < 4 + ZO && z < nz;
    // the last tile along an axis also writes the grid's last cells that lie in its upper halo
    // (lanes 61..63, rows 4 + YO ..): these are annulus cells (fill_fused counts the tiles so),
    // whose new values need no D- (L = 0 there), only their own operands
    const bool zW = zOut || (zt == A.nzt - 1 && l >= 4 + ZO && z < nz);
    const bool lastY = yt == A.nyt - 1;
    // per row: y inside the update region (uniform), output row (uniform), written row (uniform),
    // column offsets
    bool yU[R], rOut[R], rW[R], rOk[R];
    uint32_t own[R], moff[R], ooff[R];
    uint32_t own_sign = 0;
#pragma unroll
    for (int i = 0; i < R; i++) {
        const int y = yb + i;
        yU[i] = y >= 4 && y < ny - 4;
        rOut[i] = rb + i >= 4 && rb + i < 4 + YO && y < ny;
        rW[i] = rOut[i] || (lastY && rb + i >= 4 + YO && y < ny);
        bool ok;
        own[i] = col_off(y, z, ny, nz, FST, own_sign, ok);
        rOk[i] = ok;
        moff[i] = ((yU[i] && zU) || (EPI != EPI_NONE && rW[i] && zW)) ? own[i] : 0u;
        ooff[i] = (rW[i] && zW) ? own[i] : 0u;
    }
    bool anyOut = false, anyW = false;  // the wave holds an output (a written) row (uniform)
#pragma unroll
    for (int i = 0; i < R; i++) {
        anyOut = anyOut || rOut[i];
        anyW = anyW || rW[i];
    }
    // border cells of the staged plane owned by this thread
    uint32_t boff[NBT], bsign[NBT];
    bool b_ok[NBT], has_b[NBT];
    int blds[NBT];
#pragma unroll
    for (int k = 0; k < NBT; k++) {
        const int bi = t + k * WIDE_THREADS;
        has_b[k] = bi < NB;
        boff[k] = 0;
        bsign[k] = 0;
        b_ok[k] = false;
        blds[k] = 0;
        if (has_b[k]) {
            int sy, sz;
            if (bi < 3 * LZ) {
                sy = bi / LZ;
                sz = bi - sy * LZ;
            } else if (bi < 3 * LZ + 7 * NROW) {
                const int m = bi - 3 * LZ, c = m % 7;
                sy = 3 + m / 7;
                sz = c < 3 ? c : ZW + c;
            } else {
                const int m = bi - 3 * LZ - 7 * NROW;
                sy = 3 + NROW + m / LZ;
                sz = m - (sy - 3 - NROW) * LZ;
            }
            boff[k] = col_off(y0 - 7 + sy, z0 - 7 + sz, ny, nz, FST, bsign[k], b_ok[k]);
            blds[k] = sy * LZ + sz;
        }
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
    constexpr bool QT = PHYS != TTI;
    const int kz_s = g.fs ? nz - 1 - z : min(z, nz - 1 - z);
    auto ld = [&](const float *a, int xx, uint32_t off) -> float { return gld(plane(a, xx, nx, nynz), off); };
    auto opnd = [&](int k, int xx, int i, uint32_t off) -> float {
        if (QT && k == 1 && A.qtab) {
            const int y = yb + i;
            int ks = min(min(min(xx, nx - 1 - xx), min(y, ny - 1 - y)), kz_s);
            ks = ks < 0 ? 0 : (ks > A.nsponge ? A.nsponge : ks);
            return A.qtab[ks];
        }
        return ld(oin[k], xx, off);
    };
    auto fixp = [&](float v, int i, int xx) -> float {
        v = (rOk[i] && xx >= 0 && xx < nx) ? v : 0.0f;
        if constexpr (FST) v = flip(v, own_sign);
        return v;
    };
    auto fixb = [&](float v, int k, int xx) -> float {
        v = (b_ok[k] && xx >= 0 && xx < nx) ? v : 0.0f;
        if constexpr (FST) v = flip(v, bsign[k]);
        return v;
    };

    // pq[f][i][s]: p (m) of plane xx in slot (xx - xfirst) & 7; oxq likewise for the
    // x-temporaries; bq: b of the planes x .. x+3 (the update's buoyancy), ring by plane & 3
    float pq[NF][R][8], pn[NF][R], oxq[NF][R][8], bq[4][R], bn[NF][NBT];
    float mc[NM][R], oc[NO][R];
    const int xfirst = xs - 7;
#pragma unroll
    for (int fi = 0; fi < NF; fi++)
#pragma unroll
        for (int i = 0; i < R; i++) {
#pragma unroll
            for (int j = 0; j < 8; j++) oxq[fi][i][j] = 0.0f;
#pragma unroll
            for (int j = 0; j < 7; j++)
                pq[fi][i][j] = fixp(ld(fld[fi], xfirst + j, own[i]), i, xfirst + j);
            pq[fi][i][7] = 0.0f;
            pn[fi][i] = ld(fld[fi], xfirst + 7, own[i]);
        }
#pragma unroll
    for (int fi = 0; fi < NF; fi++)
#pragma unroll
        for (int k = 0; k < NBT; k++) bn[fi][k] = ld(fld[fi], xfirst + 3, boff[k]);
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < R; i++) bq[j][i] = 0.0f;
#pragma unroll
    for (int k = 0; k < NM; k++)
#pragma unroll
        for (int i = 0; i < R; i++) mc[k][i] = ld(mat[k], xfirst + 3, moff[i]);
#pragma unroll
    for (int k = 0; k < NO; k++)
#pragma unroll
        for (int i = 0; i < R; i++) oc[k][i] = opnd(k, xfirst, i, ooff[i]);

    // LDS lane bases, one per array, with this thread's first row folded in: every access below
    // is a base plus a compile-time offset (each array is under the 64 KiB reach of a ds offset)
    constexpr int SPP = (ZW + 7) * (NROW + 7);  // one staged plane of one field
    constexpr int SOY = (NROW + 2 * R) * ZW, SOZ = (YO + 2 * R) * ZW;
    const uint32_t sp_l0 = lds_addr(&S.sp[0][0][0]) + 4u * (uint32_t)(rb * LZ + l);
    const uint32_t soy_l0 = lds_addr(&S.soy[0][0][0]) + 4u * (uint32_t)((rb - 4 + R) * ZW + l);
    const uint32_t soz_l0 = lds_addr(&S.soz[0][0][0]) + 4u * (uint32_t)((rb - 4 + R) * ZW + l - 4);
    // free-surface tile (z = l - 4): the even mirror of D- reads lane -zz-1+4 for zz = l-8+j < 0,
    // i.e. offset 7 - j from this base
    const uint32_t smir_l0 = soz_l0 + 4u * (uint32_t)(8 - 2 * l);
    uint32_t bl0[NBT];
#pragma unroll
    for (int k = 0; k < NBT; k++) bl0[k] = lds_addr(&S.sp[0][0][0]) + 4u * (uint32_t)blds[k];

    auto iter = [&](auto U, int x) {
        constexpr int u = decltype(U)::value;  // x - xfirst = u (mod 8)
        // this step's copies of the column offsets, through an empty asm so that their zero
        // extension stays next to the accesses (global_load ... saddr: no 64-bit VGPR addresses)
        uint32_t own_s[R], moff_s[R], ooff_s[R], boff_s[NBT];
#pragma unroll
        for (int i = 0; i < R; i++) {
            own_s[i] = launder(own[i]);
            moff_s[i] = launder(moff[i]);
            ooff_s[i] = launder(ooff[i]);
        }
#pragma unroll
        for (int k = 0; k < NBT; k++) boff_s[k] = launder(boff[k]);
        uint32_t sp_l = sp_l0, soy_l = soy_l0, soz_l = soz_l0, smir_l = smir_l0, bl[NBT];
        asm volatile("" : "+v"(sp_l), "+v"(soy_l), "+v"(soz_l));
        if constexpr (FST) asm volatile("" : "+v"(smir_l));
#pragma unroll
        for (int k = 0; k < NBT; k++) {
            bl[k] = bl0[k];
            asm volatile("" : "+v"(bl[k]));
        }
        // staged plane: row k of this thread's window (k = 0 is tile row rb), column c = z - z0 + 7
        auto SP = [&](int buf, int fi, int k, int c) -> lds_float & {
            return lds_at(sp_l)[(buf * NF + fi) * SPP + k * LZ + c];
        };
        // y-temporaries: ring slot, row rb - 4 + k; z-temporaries: row rb - 4 + i, lane l - 4 + j
        auto OY = [&](int slot, int fi, int k) -> lds_float & {
            return lds_at(soy_l)[(slot * NF + fi) * SOY + k * ZW];
        };
        auto OZ = [&](int slot, int fi, int i, int j) -> lds_float & {
            return lds_at(soz_l)[(slot * NF + fi) * SOZ + i * ZW + j];
        };
        // queue slot of p(x + j), and of the x-temporary of plane x - 4 + j
        auto qs = [](int j) constexpr { return (u + j) & 7; };
        auto qx = [](int j) constexpr { return (u + 4 + j) & 7; };
        // the plane index through an empty asm: the per-plane addresses are formed here, each
        // step, instead of being precomputed for all eight steps (which exhausts the SGPRs)
        asm volatile("" : "+s"(x));
        constexpr int s2 = (u + 3) & 1;        // p-plane buffer of plane x+3
        constexpr int sw = (u + 3) & 3;        // ring slot of plane x+3 (written now)
        constexpr int sr = u & 3;              // ring slot of plane x (read now)
        // ---- 1. p(x+7) into its slot, stage plane x+3 in LDS ---------------------------
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
#pragma unroll
            for (int i = 0; i < R; i++) {
                pq[fi][i][qs(7)] = fixp(pn[fi][i], i, x + 7);
                SP(s2, fi, i + 3, 3) = pq[fi][i][qs(3)];
            }
#pragma unroll
            for (int k = 0; k < NBT; k++)
                if (has_b[k]) lds_at(bl[k])[(s2 * NF + fi) * SPP] = fixb(bn[fi][k], k, x + 3);
        }
        // ---- 2. request p(x+8) and the border of plane x+4 -------------------------------
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
#pragma unroll
            for (int i = 0; i < R; i++) pn[fi][i] = ld(fld[fi], x + 8, own_s[i]);
#pragma unroll
            for (int k = 0; k < NBT; k++) bn[fi][k] = ld(fld[fi], x + 4, boff_s[k]);
        }
        __syncthreads();
        WIDE_FENCE_OFF();
        // ---- 3. gradient + material sandwich at (x+3, rows, z) --------------------------
        // every row is evaluated (straight-line code the scheduler can interleave); cells
        // outside the update region take 0, as in the reference's zeroed annulus
        const int xp = x + 3;
        const bool xpU = xp >= 4 && xp < nx - 4;
        float win[NF][R + 7];  // staged rows rb .. rb+R+6 of this lane (D+y operands)
#pragma unroll
        for (int fi = 0; fi < NF; fi++)
#pragma unroll
            for (int k = 0; k < R + 7; k++)
                win[fi][k] = (k >= 3 && k < 3 + R) ? pq[fi][k - 3][qs(3)] : float(SP(s2, fi, k, 3));
        float ox[R][NF], oy[R][NF], oz[R][NF];
#pragma unroll
        for (int i = 0; i < R; i++) {
            float dx[NF], dy[NF], dz[NF];
#pragma unroll
            for (int fi = 0; fi < NF; fi++) {
                float a[8];
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = win[fi][i + j];  // rows y-3 .. y+4
                dy[fi] = g.idy * sten8(a);
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = j == 3 ? win[fi][i + 3] : float(SP(s2, fi, i + 3, j));  // z-3 .. z+4
                dz[fi] = g.idz * sten8(a);
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = pq[fi][i][qs(j)];  // planes x .. x+7
                dx[fi] = g.idx * sten8(a);
                if (FST) {  // free-surface row: x/y derivatives vanish
                    dx[fi] = z == 0 ? 0.0f : dx[fi];
                    dy[fi] = z == 0 ? 0.0f : dy[fi];
                }
            }
            if constexpr (PHYS == ISO) {
#pragma unroll
                for (int fi = 0; fi < NF; fi++) {
                    ox[i][fi] = mc[0][i] * dx[fi];
                    oy[i][fi] = mc[0][i] * dy[fi];
                    oz[i][fi] = mc[0][i] * dz[fi];
                }
            } else if constexpr (PHYS == VTI) {
                vti_material(dx[0], dy[0], dz[0], dx[1], dy[1], dz[1], mc[1][i], mc[2][i], mc[3][i], mc[0][i],
                             ox[i][0], oy[i][0], oz[i][0], ox[i][1], oy[i][1], oz[i][1]);
            } else {
                tti_material(dx[0], dy[0], dz[0], dx[1], dy[1], dz[1], mc[4][i], mc[5][i], mc[6][i], mc[7][i],
                             mc[1][i], mc[2][i], mc[3][i], mc[0][i], ox[i][0], oy[i][0], oz[i][0], ox[i][1],
                             oy[i][1], oz[i][1]);
            }
            const bool c = xpU && yU[i] && zU;
#pragma unroll
            for (int fi = 0; fi < NF; fi++) {
                ox[i][fi] = c ? ox[i][fi] : 0.0f;
                oy[i][fi] = c ? oy[i][fi] : 0.0f;
                oz[i][fi] = c ? oz[i][fi] : 0.0f;
            }
        }
#pragma unroll
        for (int i = 0; i < R; i++) {
#pragma unroll
            for (int fi = 0; fi < NF; fi++) {
                oxq[fi][i][qx(7)] = ox[i][fi];
                OY(sw, fi, i + 4) = oy[i][fi];
            }
            bq[sw][i] = mc[0][i];
        }
#pragma unroll
        for (int i = 0; i < R; i++)
            if (rb + i >= 4 && rb + i < 4 + YO) {
#pragma unroll
                for (int fi = 0; fi < NF; fi++) OZ(sw, fi, i, 4) = oz[i][fi];
            }
        // ---- 4. request the material of plane x+4 (its registers are free now) ----------
#pragma unroll
        for (int k = 0; k < NM; k++)
#pragma unroll
            for (int i = 0; i < R; i++) mc[k][i] = ld(mat[k], x + 4, moff_s[i]);
        WIDE_FENCE_OFF();
        // ---- 5. D- and leapfrog at plane x (temporaries of x written three steps ago) ----
        const bool xOut = x >= xs && x < xe;
        if (xOut && anyW && zW) {
            const bool xU = x >= 4 && x < nx - 4;
            // L of every row (the rings are padded, so a non-output row of a wave with output
            // rows reads its neighbours' slots harmlessly); inside the region D-x + D-y + D-z,
            // else 0.  Annulus-only lanes and waves (the last tile's upper halo) read no ring.
            float Lr[R][NF];
#pragma unroll
            for (int i = 0; i < R; i++)
#pragma unroll
                for (int fi = 0; fi < NF; fi++) Lr[i][fi] = 0.0f;
            if (anyOut && zOut) {
                float ywin[NF][R + 7];  // ring rows rb-4 .. rb+R+2 of this lane (D-y operands)
#pragma unroll
                for (int fi = 0; fi < NF; fi++)
#pragma unroll
                    for (int k = 0; k < R + 7; k++) ywin[fi][k] = OY(sr, fi, k);
#pragma unroll
                for (int i = 0; i < R; i++) {
                    const bool upd = xU && yU[i] && zU && !(FST && z == 0);
#pragma unroll
                    for (int fi = 0; fi < NF; fi++) {
                        float a[8];
#pragma unroll
                        for (int j = 0; j < 8; j++) a[j] = oxq[fi][i][qx(j)];  // planes x-4 .. x+3
                        const float dxm = g.idx * sten8(a);
#pragma unroll
                        for (int j = 0; j < 8; j++) a[j] = ywin[fi][i + j];  // rows y-4 .. y+3
                        const float dym = g.idy * sten8(a);
                        if (FST) {
#pragma unroll
                            for (int j = 0; j < 8; j++) {
                                // even mirror: t[-j] = t[j-1]; kz = zz < 0 reads kz' = -zz-1
                                const int zz = z - 4 + j;
                                const uint32_t o = 4u * (uint32_t)((sr * NF + fi) * SOZ + i * ZW);
                                a[j] = *lds_at(zz >= 0 ? soz_l + o + 4u * j : smir_l + o + 4u * (7 - j));
                            }
                        } else {
#pragma unroll
                            for (int j = 0; j < 8; j++) a[j] = OZ(sr, fi, i, j);  // z-4 .. z+3
                        }
                        const float dzm = g.idz * sten8(a);
                        Lr[i][fi] = upd ? dxm + dym + dzm : 0.0f;
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < R; i++) {
                if (!rW[i]) continue;
                const uint32_t so = ooff_s[i];
                const bool upd = xU && yU[i] && zU;
                float L[NF];
#pragma unroll
                for (int fi = 0; fi < NF; fi++) L[fi] = Lr[i][fi];
                const float V = oc[0][i];
                float pn_[NF];
                if (upd) {
                    const float dt2V2_B = A.dt2 * V * V / bq[sr][i];
                    const float qk = oc[1][i];
                    pn_[0] = leapfrog(dt2V2_B, L[0], qk, pq[0][i][qs(0)], oc[2][i]);
                    if constexpr (NF == 2) pn_[1] = leapfrog(dt2V2_B, L[1], qk, pq[1][i][qs(0)], oc[3][i]);
                } else {
#pragma unroll
                    for (int fi = 0; fi < NF; fi++) pn_[fi] = oc[2 + fi][i];
                }
                if constexpr (EPI == EPI_NONE) {
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
                    const float sc = V * V / bq[sr][i];
                    gst(L[0] * sc, A.e_dp, x, nynz, so);
                    if constexpr (NF == 2) gst(L[1] * sc, A.e_dm, x, nynz, so);
                } else if constexpr (EPI == EPI_BORN_INJ) {
                    const float B = bq[sr][i], dv = oc[IB][i];
                    if constexpr (PHYS == ISO) {
                        const float dt2v2OverB = A.dt * A.dt * V * V / B;
                        const float factorV = 2 * B * dv / (V * V * V);
                        pn_[0] += dt2v2OverB * oc[IB + 1][i] * factorV;
                    } else {
                        const float factor = 2 * A.dt * A.dt * dv / V;
                        pn_[0] += factor * oc[IB + 1][i];
                        pn_[1] += factor * oc[IB + 2][i];
                    }
                    gst(pn_[0], A.pOld, x, nynz, so);
                    if constexpr (NF == 2) gst(pn_[1], A.mOld, x, nynz, so);
                } else if constexpr (EPI == EPI_BORN_COUPLED) {
                    if (upd) gst(pn_[0], A.pOld, x, nynz, so);
                    const float B = bq[sr][i], dv = oc[IB][i];
                    const float sc = V * V / B;
                    const float DP = L[0] * sc;
                    gst(DP, A.e_dp, x, nynz, so);
                    const float dt2v2OverB = A.dt * A.dt * V * V / B;
                    const float factorV = 2 * B * dv / (V * V * V);
                    pn_[1] += dt2v2OverB * DP * factorV;
                    gst(pn_[1], A.mOld, x, nynz, so);
                } else {  // EPI_IMAGE
                    if (upd) {
                        gst(pn_[0], A.pOld, x, nynz, so);
                        if constexpr (NF == 2) gst(pn_[1], A.mOld, x, nynz, so);
                    }
                    const float B = bq[sr][i];
                    float gk = oc[IB][i];
                    if constexpr (PHYS == ISO) {
                        const float factorV = +2 * B / (V * V * V);
                        gk += factorV * oc[IB + 1][i] * pq[0][i][qs(0)];
                    } else {
                        const float factor = 2 * B / (V * V * V);
                        gk += factor * (oc[IB + 1][i] * pq[0][i][qs(0)] + oc[IB + 2][i] * pq[1][i][qs(0)]);
                    }
                    gst(gk, A.e_dp, x, nynz, so);
                }
                WIDE_FENCE_OFF();
            }
        }
        WIDE_FENCE_OFF();
        // ---- 6. request the output operands of plane x+1 --------------------------------
#pragma unroll
        for (int k = 0; k < NO; k++)
#pragma unroll
            for (int i = 0; i < R; i++) oc[k][i] = opnd(k, x + 1, i, ooff_s[i]);
    };

    // steps x < xs only fill the queues; steps x >= xe (at most 7, to keep the 8-step unrolling)
    // store nothing
    for (int xb = xfirst; xb < xe; xb += 8) {
        iter(IC<0>{}, xb);
        iter(IC<1>{}, xb + 1);
        iter(IC<2>{}, xb + 2);
        iter(IC<3>{}, xb + 3);
        iter(IC<4>{}, xb + 4);
        iter(IC<5>{}, xb + 5);
        iter(IC<6>{}, xb + 6);
        iter(IC<7>{}, xb + 7);
    }
}

template <int PHYS, int EPI, int R>
__global__ void __launch_bounds__(WIDE_THREADS, wide_wpe(fused_nf(PHYS, EPI), R)) k_wide_step(FusedArgs<PHYS> A) {
    __shared__ WideLds<fused_nf(PHYS, EPI), R> S;
    const int bid = (int)blockIdx.x, xcd = bid & 7;
    int loc, zt, yt, xc;
    if (!fused_tile_of_wg(A, bid, loc, zt, yt, xc)) return;
    const int per = A.per_xcd;
    if (per > 0 && loc >= per) {
        // rounds in step, as k_fused_step: wait (bounded) until the XCD finished every earlier round
        // but `slack` workgroups
        if (threadIdx.x == 0) {
            const int t0 = loc / per * per - A.slack;
            const unsigned target = (unsigned)(t0 > 0 ? t0 : 0);
            for (int i = 0; i < 4096; i++) {
                if (__hip_atomic_load(&A.sync[xcd], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
                __builtin_amdgcn_s_sleep(16);
            }
        }
        __syncthreads();
    }
    if (A.g.fs && zt == 0)
        wide_body<PHYS, R, EPI, true>(A, S, zt, yt, xc);
    else
        wide_body<PHYS, R, EPI, false>(A, S, zt, yt, xc);
    if (per > 0) {
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_fetch_add(&A.sync[xcd], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// x-planes per chunk of the wide tile.  Where the plain VTI / TTI step keeps rounds in step
// (fused_lockstep_of, from WIDE_LOCKSTEP_ROUNDS rounds), chunks of at least 64 planes (TTI) or 96
// (VTI, whose cheaper planes leave the warm-up a larger share) balanced as
// fused_cx balances them (whole rounds against the warm-up each chunk pays) among the splits that
// give those rounds: each chunk re-synchronises the workgroups of a round, so neighbouring tiles
// keep reading their shared lines while these are in the XCD's L2.  Measured with the XCDs taking
// turns (one process): TTI 800^3 5 chunks of 160 planes (12.9 rounds) 10.25-10.28 ms, 7 of 115
// 10.21, 8 of 100 10.38, 6 of 134 10.48, 4 of 200 10.73; VTI 800^3 7.88, 7.94, 8.10, 8.16 and
// 8.35 ms -- the order of that balance.  Where 64-plane chunks alone give the rounds: TTI 640^3
// 5.83-5.92 ms in step against 5.94-6.06 for 128 planes without, VTI 640^3 4.59 against 4.40-4.42.
// Otherwise fused_cx without rounds in step.
constexpr long WIDE_LOCKSTEP_ROUNDS = 12;
static inline int wide_cx(const Prop *p, long tiles, int phys, int epi) {
    if (p->fused_chunk > 0 || std::getenv("WAVEFD_FUSED_CX")) return fused_cx(p, tiles, 8);
    const long cus = p->cu_count > 0 ? p->cu_count : 256;
    const bool auto_ls = p->fused_lockstep < 0 && !std::getenv("WAVEFD_FUSED_LOCKSTEP");
    if (phys != ISO && epi == EPI_NONE && cus % 8 == 0 && (auto_ls || p->fused_lockstep == 1)) {
        const long min_cx = phys == TTI ? 64 : 96;
        long best_cx = 0;
        double best = 1e300;
        for (long nxc = 1; nxc <= 64 && nxc <= p->nx; nxc++) {
            const long cx = (p->nx + nxc - 1) / nxc;
            if (cx < min_cx) break;
            if ((p->nx + cx - 1) / cx != nxc || tiles * nxc < WIDE_LOCKSTEP_ROUNDS * cus) continue;
            const double cost = (double)((tiles * nxc + cus - 1) / cus) * (double)chunk_steps(p->nx, nxc, 8);
            if (cost < best * 0.999) {
                best = cost;
                best_cx = cx;
            }
        }
        if (best_cx > 0) return (int)best_cx;
    }
    return fused_cx(p, tiles, 8);
}

// the same rule for two resident workgroups per CU
static inline int wide_cx2(const Prop *p, long tiles, int, int) { return fused_cx(p, tiles, 8, 2); }

template <int PHYS, int EPI, int R = wide_r(PHYS, EPI)>
static inline bool launch_wide(Prop *p, bool linear, float *e0, float *e1, float *e2) {
    FusedArgs<PHYS> A;
    constexpr bool two_per_cu = wide_wpe(fused_nf(PHYS, EPI), R) == 4;
    const long ntiles =
        fill_fused<PHYS, 64, 8 * R, EPI>(p, linear, e0, e1, e2, A, 8, two_per_cu ? wide_cx2 : wide_cx);
    const bool lockstep = fused_lockstep_of(p, PHYS, EPI, ntiles, WIDE_LOCKSTEP_ROUNDS);
    if (lockstep && p->fused_sync) {
        A.per_xcd = p->cu_count / 8;
        // a quarter of a round of slack (with contiguous runs per XCD: TTI 800^3 11.02 against 11.09
        // ms for the strict barrier, VTI 800^3 8.39 against 8.51 ms; with the XCDs taking turns
        // 10.40 against 10.44 ms)
        if (A.slack < 0) A.slack = A.per_xcd / 4;
        k_zero_u32<<<1, 64, 0, p->stream>>>(p->fused_sync, 8);
        WFD_LAUNCHED();
    }
    // the XCDs take turns: group i of G tiles (the workgroups one round of an XCD holds) goes to
    // XCD i mod 8, so the eight XCDs run neighbouring groups at the same time and share their halo
    // lines through the memory-side cache, where contiguous runs per XCD put them chunks apart
    // (TTI 800^3 10.40 against 11.05 ms, VTI 512^3 2.28 against 2.39, TTI 512^3 3.17 against 3.21,
    // Iso 512^3 1.06 against 1.09, Iso 256^3 unchanged; WAVEFD_FUSED_XMAP=0 keeps the runs, for
    // measurements).  The grid covers the deals: XCD 0 holds the most workgroups, those of every
    // full deal and of the last one's first group; ids past the last tile do nothing.
    long grid = ntiles;
    const char *xm = std::getenv("WAVEFD_FUSED_XMAP");
    if (p->cu_count % 8 == 0 && !(xm && std::atoi(xm) == 0)) {
        const long G = (long)(p->cu_count / 8) * (two_per_cu ? 2 : 1), deals = ntiles / (8 * G);
        A.xmap = (int)G;
        grid = 8 * (deals * G + std::min(G, ntiles - deals * 8 * G));
    }
    k_wide_step<PHYS, EPI, R><<<dim3((unsigned)grid), dim3(WIDE_THREADS), 0, p->stream>>>(A);
    return true;
}

// Several grids stepped by one launch of the wide kernel: grid s owns workgroups [wg[s], wg[s+1]),
// each range a multiple of 8 long (the XCD-aware numbering holds inside every range), with its
// own sizes, arrays and chunking (each grid's own rule); no rounds in step.  The grids share the
// physics, the rows per thread and the epilogue.
template <int PHYS, int EPI, int R>
__global__ void __launch_bounds__(WIDE_THREADS, wide_wpe(fused_nf(PHYS, EPI), R)) k_wide_batch(FusedBatch<PHYS> B) {
    __shared__ WideLds<fused_nf(PHYS, EPI), R> S;
    const int bid = (int)blockIdx.x;
    int s = 0;
    while (s + 1 < B.n && bid >= B.wg[s + 1]) s++;
    const FusedArgs<PHYS> &A = B.a[s];
    int loc, zt, yt, xc;
    if (!fused_tile_of_wg(A, bid - B.wg[s], loc, zt, yt, xc)) return;
    if (A.g.fs && zt == 0)
        wide_body<PHYS, R, EPI, true>(A, S, zt, yt, xc);
    else
        wide_body<PHYS, R, EPI, false>(A, S, zt, yt, xc);
}

template <int PHYS, int EPI, int R = wide_r(PHYS, EPI)>
static inline bool launch_wide_batch(Prop **ps, int n, float *const *e0, float *const *e1, float *const *e2) {
    FusedBatch<PHYS> B = {};
    B.n = n;
    constexpr bool two_per_cu = wide_wpe(fused_nf(PHYS, EPI), R) == 4;
    long wg = 0;
    for (int i = 0; i < n; i++) {
        B.wg[i] = (int)wg;
        const long t = fill_fused<PHYS, 64, 8 * R, EPI>(ps[i], false, e0 ? e0[i] : nullptr, e1 ? e1[i] : nullptr,
                                                       e2 ? e2[i] : nullptr, B.a[i], 8,
                                                       two_per_cu ? wide_cx2 : wide_cx);
        wg += (t + 7) / 8 * 8;
        if (wg >= (1L << 31)) {
            set_error("batched step: too many workgroups");
            return false;
        }
    }
    B.wg[n] = (int)wg;
    k_wide_batch<PHYS, EPI, R><<<dim3((unsigned)wg), dim3(WIDE_THREADS), 0, ps[0]->stream>>>(B);
    return true;
}

}  // namespace wfd
