// wfd_2d_fused.hpp -- single-pass fused 2D time step (D+ sandwich, D-, leapfrog with Q) for the
// 2D Iso, VTI and TTI propagators (src/prop2DAco{Iso,VTI,TTI}DenQ_DEO2_FDTD.h timeStep): one
// kernel replaces the two passes, so the temporaries never reach HBM.
//
// A workgroup of F2W threads owns a strip of F2W consecutive z (one thread per z, lanes
// coalesced), whose inner F2O = F2W - 7 cells are its outputs (D- reads temporaries at -4..+3),
// and marches over a chunk of x.  Per thread: p(x..x+7) of its column (the D+x operand) and the
// x-temporaries of planes x-4..x+3 (the D-x operand) in registers.  At step x the strip of
// p(x+3) is staged in LDS with a 7-cell border, the sandwich at (x+3, z) is formed (D+z from
// LDS), its x-temporary pushed into the queue and its z-temporary written into a 4-plane LDS
// ring; then the output threads apply D- and the update at plane x.  There is no second
// transverse axis, so the halo is 7 cells of 256 and the step moves the compulsory bytes plus
// the strips' borders.
//
// TTI: the weighted differences D-x(cos g1), D-z(sin g1), D-x(sin g3), D-z(cos g3) read products
// of a weight and a temporary at neighbouring cells; each product is formed once, when its
// temporary is (the reference's cosTheta[i] * inPG1[i] at the same i), and queued or stored in
// the ring instead of the temporary.
//
// Semantics match the two-pass kernels (wfd_2d.hpp) exactly: temporaries are 0 outside the
// update region, the free surface uses the odd mirror for D+ and the even mirror for D-, the x
// derivative vanishes at kz = 0 and L = 0 there, expression order unchanged.
//
// Epilogues (the Born loops' per-step calls, as in the 3D kernel): SAVE_DP writes DP = L v^2/b
// (DM) instead of pSpace (ScaleSpatialDerivatives); BORN_INJ adds the V source 2 dt^2 dV/v DP
// (DM) to the new field (src/prop2DAcoIsoDenQ_DEO2_FDTD.h ForwardBornInjection_V, and the VTI/TTI
// twins); IMAGE adds 2b/v^3 (DP p + DM m) of this step's input fields to g
// (AdjointBornAccumulation_V).  Every cell gets the pointwise kernel's expression, annulus
// included, where L = 0 and the new field keeps pOld's value.
#pragma once
#include "wfd_fused.hpp"
#include "wfd_kernels.hpp"
#include "wfd_runtime.hpp"

namespace wfd {

constexpr int F2W = 256;      // threads (lanes along z) per workgroup
constexpr int F2O = F2W - 7;  // output cells along z per workgroup

template <int PHYS>
struct Fused2Args {
    Grid g;
    float dt2;
    int cx, nzt, nxc;
    const float *pCur, *mCur;
    float *pOld, *mOld, *pSpace, *mSpace;
    const float *v, *b, *q, *eps, *eta, *f, *sT, *cT;
    // epilogue operands: SAVE_DP: DP, DM out (e_dp, e_dm); BORN_INJ: dV, DP, DM in (e_dv,
    // e_ip, e_im); IMAGE: g in/out (e_dp), DP, DM in (e_ip, e_im)
    float dt;
    float *e_dp, *e_dm;
    const float *e_dv, *e_ip, *e_im;
    // Q by sponge distance (nsponge + 1 values) while Q is the setup profile, else null: the
    // update then takes qtab[min(distance, nsponge)] instead of loading q
    const float *qtab;
    int nsponge;
};

template <int PHYS>
struct Fused2Lds {
    static constexpr int NF = PHYS == ISO ? 1 : 2;  // wavefields
    static constexpr int NZ = PHYS == TTI ? 2 : 1;  // z-direction operands per field
    float sp[2][NF][F2W + 7];                       // p (m) of plane x+3, cells z0-7 .. z0+F2W-1
    float sz[4][NF][NZ][F2W];                       // z operands of planes x .. x+3 (ring)
};

// the 2D material sandwich (src/prop2DAcoVTIDenQ_DEO2_FDTD.h:785-793,
// src/prop2DAcoTTIDenQ_DEO2_FDTD.h:775-783 on the rotated derivatives); the 2D
// text multiplies the coupling coefficient left to right, B * F * A * sqrt(1 - A * A)
__device__ __forceinline__ void aniso2_material(float dPx, float dPz, float dMx, float dMz, float eps, float eta,
                                                float f, float b, float &tPx, float &tPz, float &tMx, float &tMz) {
    const float E = 1 + 2 * eps;
    const float A = eta;
    const float F = f;
    const float B = b;
    tPx = B * E * dPx;
    tPz = B * (1 - F * A * A) * dPz + B * F * A * sqrtf(1 - A * A) * dMz;
    tMx = B * (1 - F) * dMx;
    tMz = B * F * A * sqrtf(1 - A * A) * dPz + B * (1 - F + F * A * A) * dMz;
}

__device__ __forceinline__ float sten8v(const float *a) {
    return C1 * (-a[3] + a[4]) + C2 * (-a[2] + a[5]) + C3 * (-a[1] + a[6]) + C4 * (-a[0] + a[7]);
}

template <int PHYS, int EPI, bool FST>
__device__ __forceinline__ void fused2_body(const Fused2Args<PHYS> &A, Fused2Lds<PHYS> &S, int zt, int xc) {
    constexpr int NF = Fused2Lds<PHYS>::NF;
    constexpr int NQ = PHYS == TTI ? 2 : 1;  // x-queues per field
    constexpr int NM = PHYS == ISO ? 1 : (PHYS == VTI ? 4 : 6);
    // output operands: v, q, pOld (, mOld) (b comes with the sandwich's material), then for the
    // epilogues (SAVE_DP) b, (BORN_INJ) dV, DP (, DM) or (IMAGE) b, g, DP (, DM)
    constexpr int NEX = EPI == EPI_NONE ? 0 : (EPI == EPI_SAVE_DP ? 1 : (EPI == EPI_BORN_INJ ? 1 + NF : 2 + NF));
    constexpr int IE = 2 + NF;  // first epilogue operand
    constexpr int NO = IE + NEX;
    static_assert(EPI == EPI_NONE || EPI == EPI_SAVE_DP || EPI == EPI_BORN_INJ || EPI == EPI_IMAGE,
                  "2D epilogues: none, save DP, Born injection V, imaging V");
    const Grid &g = A.g;
    const int nx = (int)g.nx, nz = (int)g.nz;
    const int t = threadIdx.x;
    const int z0 = zt * F2O, z = z0 - 4 + t;
    const int xs = xc * A.cx;
    const int xe = xs + A.cx < nx ? xs + A.cx : nx;
    const int zlo = FST ? 0 : 4;
    const bool zU = z >= zlo && z < nz - 4;
    const bool out_thread = t >= 4 && t < 4 + F2O && z < nz;
    // a column zz of the strip: its offset in a row (clamped into the grid), the odd mirror's
    // sign above a free surface, zero outside
    auto col = [&](int zz, uint32_t &sign, bool &ok) -> int {
        sign = 0;
        ok = false;
        if (zz >= 0 && zz < nz) {
            ok = true;
            return zz;
        }
        if (FST && zz < 0 && -zz < nz) {
            sign = 0x80000000u;
            ok = true;
            return -zz;
        }
        return 0;
    };
    uint32_t own_sign, b_sign = 0;
    bool own_ok, b_ok = false;
    const int own = col(z, own_sign, own_ok);
    const int moff = zU ? own : 0;
    const int ooff = out_thread ? own : 0;
    // border cells of the staged strip: entries 0..2 (z0-7 .. z0-5) and F2W+3 .. F2W+6
    const bool has_b = t < 7;
    const int bent = t < 3 ? t : F2W + t;
    int boff = 0;
    if (has_b) boff = col(z0 - 7 + bent, b_sign, b_ok);
    const float *fld[2] = {A.pCur, A.mCur};
    const float *mat[6] = {A.b, A.eps, A.eta, A.f, A.sT, A.cT};
    const float *oin[NO];
    oin[0] = A.v;
    oin[1] = A.q;
    oin[2] = A.pOld;
    if constexpr (NF == 2) oin[3] = A.mOld;
    if constexpr (EPI == EPI_SAVE_DP) oin[IE] = A.b;
    if constexpr (EPI == EPI_BORN_INJ) {
        oin[IE] = A.e_dv;
        oin[IE + 1] = A.e_ip;
        if constexpr (NF == 2) oin[IE + 2] = A.e_im;
    }
    if constexpr (EPI == EPI_IMAGE) {
        oin[IE] = A.b;
        oin[IE + 1] = A.e_dp;
        oin[IE + 2] = A.e_ip;
        if constexpr (NF == 2) oin[IE + 3] = A.e_im;
    }
    const long nzl = nz;
    auto ld = [&](const float *a, int xx, int off) -> float {
        const int xcl = xx < 0 ? 0 : (xx >= nx ? nx - 1 : xx);
        return a[(long)xcl * nzl + off];
    };
    // output operand k of plane xx: Q from the table while it is the profile (the index clamped
    // into the table for cells whose value is never used)
    // (the grid's free surface, not the tile's: every z-strip of a free-surface grid measures the
    // distance to the bottom row only, src/absorb.h:63)
    const int kzs = A.g.fs ? nz - 1 - z : min(z, nz - 1 - z);
    auto opnd = [&](int k, int xx, int off) -> float {
        if (k == 1 && A.qtab) {
            int ks = min(min(xx, nx - 1 - xx), kzs);
            ks = ks < 0 ? 0 : (ks > A.nsponge ? A.nsponge : ks);
            return A.qtab[ks];
        }
        return ld(oin[k], xx, off);
    };
    auto fixp = [&](float v, int xx) -> float {
        v = (own_ok && xx >= 0 && xx < nx) ? v : 0.0f;
        if constexpr (FST) v = flip(v, own_sign);
        return v;
    };
    auto fixb = [&](float v, int xx) -> float {
        v = (b_ok && xx >= 0 && xx < nx) ? v : 0.0f;
        if constexpr (FST) v = flip(v, b_sign);
        return v;
    };

    float pq[NF][8], pn[NF], bn[NF], oxq[NF][NQ][8], bq[4];
    float mc[NM], mnx[NM], oc[NO], onx[NO];
    const int xfirst = xs - 7;
#pragma unroll
    for (int fi = 0; fi < NF; fi++) {
#pragma unroll
        for (int k = 0; k < NQ; k++)
#pragma unroll
            for (int j = 0; j < 8; j++) oxq[fi][k][j] = 0.0f;
        pq[fi][0] = 0.0f;
#pragma unroll
        for (int j = 1; j < 8; j++) pq[fi][j] = fixp(ld(fld[fi], xfirst + j - 1, own), xfirst + j - 1);
        pn[fi] = ld(fld[fi], xfirst + 7, own);
        bn[fi] = ld(fld[fi], xfirst + 3, boff);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) bq[j] = 0.0f;
#pragma unroll
    for (int k = 0; k < NM; k++) mnx[k] = ld(mat[k], xfirst + 3, moff);
#pragma unroll
    for (int k = 0; k < NO; k++) onx[k] = opnd(k, xfirst, ooff);

    auto iter = [&](auto U, int x) {
        constexpr int u = decltype(U)::value;  // x - xfirst = u (mod 4)
        constexpr int s2 = (u + 3) & 1;        // strip buffer of plane x+3
        constexpr int sw = (u + 3) & 3;        // ring slot of plane x+3 (written now)
        constexpr int sr = u & 3;              // ring slot of plane x (read now)
        // ---- 1. shift the queues, stage plane x+3 ----------------------------------------
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
#pragma unroll
            for (int j = 0; j < 7; j++) pq[fi][j] = pq[fi][j + 1];
            pq[fi][7] = fixp(pn[fi], x + 7);
            S.sp[s2][fi][t + 3] = pq[fi][3];
            if (has_b) S.sp[s2][fi][bent] = fixb(bn[fi], x + 3);
        }
#pragma unroll
        for (int k = 0; k < NM; k++) mc[k] = mnx[k];
#pragma unroll
        for (int k = 0; k < NO; k++) oc[k] = onx[k];
        // ---- 2. request the operands of step x+1 ---------------------------------------
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
            pn[fi] = ld(fld[fi], x + 8, own);
            bn[fi] = ld(fld[fi], x + 4, boff);
        }
#pragma unroll
        for (int k = 0; k < NM; k++) mnx[k] = ld(mat[k], x + 4, moff);
#pragma unroll
        for (int k = 0; k < NO; k++) onx[k] = opnd(k, x + 1, ooff);
        __syncthreads();
        // ---- 3. sandwich at (x+3, z) ---------------------------------------------------
        const int xp = x + 3;
        float tq[NF][NQ], tz[NF][2];
#pragma unroll
        for (int fi = 0; fi < NF; fi++) tq[fi][0] = tq[fi][NQ - 1] = tz[fi][0] = tz[fi][1] = 0.0f;
        if (xp >= 4 && xp < nx - 4 && zU) {
            float dx[NF], dz[NF];
#pragma unroll
            for (int fi = 0; fi < NF; fi++) {
                float a[8];
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = S.sp[s2][fi][t + j];  // cells z-3 .. z+4
                dz[fi] = g.idz * sten8v(a);
                dx[fi] = (FST && z == 0) ? 0.0f : g.idx * sten8v(pq[fi]);
            }
            if constexpr (PHYS == ISO) {
                tq[0][0] = mc[0] * dx[0];
                tz[0][0] = mc[0] * dz[0];
            } else if constexpr (PHYS == VTI) {
                aniso2_material(dx[0], dz[0], dx[1], dz[1], mc[1], mc[2], mc[3], mc[0], tq[0][0], tz[0][0], tq[1][0],
                                tz[1][0]);
            } else {
                const float s = mc[4], c = mc[5];
                const float p1 = c * dx[0] - s * dz[0], p3 = s * dx[0] + c * dz[0];
                const float m1 = c * dx[1] - s * dz[1], m3 = s * dx[1] + c * dz[1];
                float t1p, t3p, t1m, t3m;
                aniso2_material(p1, p3, m1, m3, mc[1], mc[2], mc[3], mc[0], t1p, t3p, t1m, t3m);
                // operands of the weighted differences: x (queues) cos g1, sin g3; z (ring)
                // sin g1, cos g3
                tq[0][0] = c * t1p;
                tq[0][1] = s * t3p;
                tq[1][0] = c * t1m;
                tq[1][1] = s * t3m;
                tz[0][0] = s * t1p;
                tz[0][1] = c * t3p;
                tz[1][0] = s * t1m;
                tz[1][1] = c * t3m;
            }
        }
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
#pragma unroll
            for (int k = 0; k < NQ; k++) {
#pragma unroll
                for (int j = 0; j < 7; j++) oxq[fi][k][j] = oxq[fi][k][j + 1];
                oxq[fi][k][7] = tq[fi][k];
            }
#pragma unroll
            for (int k = 0; k < Fused2Lds<PHYS>::NZ; k++) S.sz[sw][fi][k][t] = tz[fi][k];
        }
        bq[sw] = mc[0];
        // ---- 4. D- and leapfrog at plane x -------------------------------------------
        if (x < xs || x >= xe || !out_thread) return;
        const bool upd = x >= 4 && x < nx - 4 && zU;
        const long k0 = (long)x * nzl + ooff;
        if constexpr (EPI == EPI_NONE) {
            if (!upd) {
                if (A.pSpace) {
                    A.pSpace[k0] = 0.0f;
                    if constexpr (NF == 2) A.mSpace[k0] = 0.0f;
                }
                return;
            }
        }
        float L[NF];
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
            if (!upd) {
                L[fi] = 0.0f;
                continue;
            }
            if (FST && z == 0) {
                L[fi] = 0.0f + 0.0f;
                continue;
            }
            // the z operands at z-4 .. z+3 (even mirror above a free surface: t[-j] = t[j-1])
            auto zcell = [&](int j) -> int {
                const int zz = z - 4 + j;
                const int zm = (FST && zz < 0) ? -zz - 1 : zz;
                return zm - (z0 - 4);
            };
            float a[8];
            if constexpr (PHYS == TTI) {
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = S.sz[sr][fi][0][zcell(j)];
                const float B1 = sten8v(a);
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = S.sz[sr][fi][1][zcell(j)];
                const float B3 = sten8v(a);
                const float A1 = sten8v(oxq[fi][0]);
                const float A3 = sten8v(oxq[fi][NQ - 1]);
                const float o1 = g.idx * A1 - g.idz * B1;
                const float o3 = g.idx * A3 + g.idz * B3;
                L[fi] = o1 + o3;
            } else {
                const float dxm = g.idx * sten8v(oxq[fi][0]);
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = S.sz[sr][fi][0][zcell(j)];
                const float dzm = g.idz * sten8v(a);
                L[fi] = dxm + dzm;
            }
        }
        const float V = oc[0];
        // the new field: updated inside the region, pOld's value in the annulus
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
            A.pOld[k0] = pn_[0];
            if (A.pSpace) A.pSpace[k0] = L[0];
            if constexpr (NF == 2) {
                A.mOld[k0] = pn_[1];
                if (A.mSpace) A.mSpace[k0] = L[1];
            }
        } else if constexpr (EPI == EPI_SAVE_DP) {
            if (upd) {
                A.pOld[k0] = pn_[0];
                if constexpr (NF == 2) A.mOld[k0] = pn_[1];
            }
            const float sc = V * V / oc[IE];  // ScaleSpatialDerivatives: pSpace *= v^2/b
            A.e_dp[k0] = L[0] * sc;
            if constexpr (NF == 2) A.e_dm[k0] = L[1] * sc;
        } else if constexpr (EPI == EPI_BORN_INJ) {
            const float factor = 2 * A.dt * A.dt * oc[IE] / V;
            pn_[0] += factor * oc[IE + 1];
            A.pOld[k0] = pn_[0];
            if constexpr (NF == 2) {
                pn_[1] += factor * oc[IE + 2];
                A.mOld[k0] = pn_[1];
            }
        } else {  // EPI_IMAGE: g += 2b/v^3 (DP p + DM m) with this step's input fields
            if (upd) {
                A.pOld[k0] = pn_[0];
                if constexpr (NF == 2) A.mOld[k0] = pn_[1];
            }
            const float B = oc[IE];
            float gk = oc[IE + 1];
            if constexpr (NF == 1) {
                const float factorV = +2 * B / (V * V * V);
                gk += factorV * oc[IE + 2] * pq[0][0];
            } else {
                const float factor = 2 * B / (V * V * V);
                gk += factor * (oc[IE + 2] * pq[0][0] + oc[IE + 3] * pq[1][0]);
            }
            A.e_dp[k0] = gk;
        }
    };

    // steps x < xs only fill the queues; steps x >= xe (at most 3) store nothing
    for (int xb = xfirst; xb < xe; xb += 4) {
        iter(IC<0>{}, xb);
        iter(IC<1>{}, xb + 1);
        iter(IC<2>{}, xb + 2);
        iter(IC<3>{}, xb + 3);
    }
}

template <int PHYS, int EPI>
__global__ void __launch_bounds__(F2W) k2_fused_step(Fused2Args<PHYS> A) {
    __shared__ Fused2Lds<PHYS> S;
    const int bid = (int)blockIdx.x;
    const int zt = bid % A.nzt, xc = bid / A.nzt;
    if (xc >= A.nxc) return;
    if (A.g.fs && zt == 0)
        fused2_body<PHYS, EPI, true>(A, S, zt, xc);
    else
        fused2_body<PHYS, EPI, false>(A, S, zt, xc);
}

// x-planes per chunk: enough workgroups for about eight per CU, chunks of at least 32 planes
// (each pays a 7-step warm-up).  Measured at 8192^2: 8 per CU against 4 per CU, VTI 1.02 ->
// 0.79-0.94 ms, TTI 1.20 -> 1.03 ms, Iso 0.470 -> 0.463 ms; 2 per CU is slower, 12 and 16 the same
static inline int fused2_cx(const Prop *p, int nzt) {
    const long want = 8L * (p->cu_count > 0 ? p->cu_count : 256);
    long nxc = (want + nzt - 1) / nzt;
    nxc = std::max(1L, std::min(nxc, p->nx / 32 > 0 ? p->nx / 32 : 1L));
    return (int)((p->nx + nxc - 1) / nxc);
}

// e0, e1, e2: the epilogue's arrays (device): SAVE_DP (DP, DM, -), BORN_INJ (dV, DP, DM),
// IMAGE (g, DP, DM); the epilogues leave pSpace/mSpace as they were
template <int PHYS, int EPI = EPI_NONE>
static inline bool launch_fused2(Prop *p, bool linear, float *e0 = nullptr, float *e1 = nullptr,
                                 float *e2 = nullptr) {
    const Field *f = p->f;
    Fused2Args<PHYS> A;
    A.g = p->g;
    A.dt2 = p->dt * p->dt;
    A.dt = p->dt;
    A.e_dp = EPI == EPI_SAVE_DP || EPI == EPI_IMAGE ? e0 : nullptr;
    A.e_dm = EPI == EPI_SAVE_DP ? e1 : nullptr;
    A.e_dv = EPI == EPI_BORN_INJ ? e0 : nullptr;
    A.e_ip = EPI == EPI_BORN_INJ || EPI == EPI_IMAGE ? e1 : nullptr;
    A.e_im = EPI == EPI_BORN_INJ || EPI == EPI_IMAGE ? e2 : nullptr;
    if (EPI != EPI_NONE) linear = true;  // pSpace/mSpace untouched
    A.nzt = (int)((p->nz + F2O - 1) / F2O);
    A.cx = fused2_cx(p, A.nzt);
    A.nxc = (int)((p->nx + A.cx - 1) / A.cx);
    A.pCur = f[F_PCUR].d;
    A.pOld = f[F_POLD].d;
    A.pSpace = linear ? nullptr : f[F_PSPACE].d;
    A.mCur = PHYS == ISO ? nullptr : f[F_MCUR].d;
    A.mOld = PHYS == ISO ? nullptr : f[F_MOLD].d;
    A.mSpace = (PHYS == ISO || linear) ? nullptr : f[F_MSPACE].d;
    A.v = f[F_V].d;
    A.b = f[F_B].d;
    A.q = f[F_Q].d;
    A.qtab = p->q_profile ? p->qtab : nullptr;
    A.nsponge = (int)p->nsponge;
    A.eps = PHYS == ISO ? nullptr : f[F_EPS].d;
    A.eta = PHYS == ISO ? nullptr : f[F_ETA].d;
    A.f = PHYS == ISO ? nullptr : f[F_F].d;
    A.sT = PHYS == TTI ? f[F_ST].d : nullptr;
    A.cT = PHYS == TTI ? f[F_CT].d : nullptr;
    if (PHYS == ISO) A.mCur = A.pCur, A.mOld = A.pOld;  // never read: one wavefield
    const long nwg = (long)A.nzt * A.nxc;
    k2_fused_step<PHYS, EPI><<<dim3((unsigned)nwg), dim3(F2W), 0, p->stream>>>(A);
    WFD_LAUNCHED();
    return true;
}

}  // namespace wfd
