// wfd_2d.hpp -- the 2D propagators (src/prop2DAco{Iso,VTI,TTI}DenQ_DEO2_FDTD.h): the same
// handle (Prop) with ny = 1 and dim = 2, two-pass operator kernels over (x, z), z fastest
// (k = kx nz + kz, the reference's layout).  One thread per cell; every kernel writes its whole
// output, annulus included, so each of the reference's zero / interior / free-surface loops
// folds into one launch:
//   * update region: kx in [4, nx-4), kz in [4, nz-4), plus kz = 0..3 under a free surface;
//   * D+ (and the sandwiches): odd mirror above a free surface, x derivative 0 at kz = 0
//     (src/propagatorStaticFunctions.h:79-161); D-: even mirror, both derivatives 0 at kz = 0
//     (:233-310); the time updates use L = 0 at kz = 0 and still update there;
//   * the mirrored operand is loaded with its sign flipped, so each expression keeps the
//     reference's terms and order (-(-a) = a exactly).
#pragma once
#include "wfd_2d_fused.hpp"
#include "wfd_kernels.hpp"
#include "wfd_runtime.hpp"

namespace wfd {

inline dim3 op2_block() { return dim3(64, 4, 1); }
inline dim3 op2_grid(const Grid &g) { return dim3((unsigned)((g.nz + 63) / 64), (unsigned)((g.nx + 3) / 4), 1); }

struct Cell2 {
    long kx, kz, k;
    bool in;    // inside the grid
    bool upd;   // inside the update region
};

__device__ __forceinline__ Cell2 cell2(const Grid &g) {
    Cell2 c;
    c.kz = (long)blockIdx.x * 64 + threadIdx.x;
    c.kx = (long)blockIdx.y * 4 + threadIdx.y;
    c.in = c.kz < g.nz && c.kx < g.nx;
    c.k = c.kx * g.nz + c.kz;
    const long zlo = g.fs ? 0 : 4;
    c.upd = c.in && c.kx >= 4 && c.kx < g.nx - 4 && c.kz >= zlo && c.kz < g.nz - 4;
    return c;
}

// a[kz + o] along z for o in -4..+4, with the free surface's odd (D+) or even (D-) mirror
// applied to indices above it (only reached from rows kz < 4 under a free surface)
__device__ __forceinline__ float zodd(const float *a, long base, long kz, long o) {
    const long zz = kz + o;
    return zz >= 0 ? a[base + zz] : -a[base - zz];
}
__device__ __forceinline__ float zeven(const float *a, long base, long kz, long o) {
    const long zz = kz + o;
    return zz >= 0 ? a[base + zz] : a[base - zz - 1];
}

// D+ along x and z at (kx, kz) (src/propagatorStaticFunctions.h:60-76)
__device__ __forceinline__ void dplus2(const Grid &g, const float *inX, const float *inZ, const Cell2 &c, float &dx,
                                       float &dz) {
    const long nz = g.nz, kx = c.kx, kz = c.kz, b = kx * nz;
    const float sx = C1 * (-inX[(kx + 0) * nz + kz] + inX[(kx + 1) * nz + kz]) +
                     C2 * (-inX[(kx - 1) * nz + kz] + inX[(kx + 2) * nz + kz]) +
                     C3 * (-inX[(kx - 2) * nz + kz] + inX[(kx + 3) * nz + kz]) +
                     C4 * (-inX[(kx - 3) * nz + kz] + inX[(kx + 4) * nz + kz]);
    const float sz = C1 * (-zodd(inZ, b, kz, 0) + zodd(inZ, b, kz, 1)) + C2 * (-zodd(inZ, b, kz, -1) + zodd(inZ, b, kz, 2)) +
                     C3 * (-zodd(inZ, b, kz, -2) + zodd(inZ, b, kz, 3)) +
                     C4 * (-zodd(inZ, b, kz, -3) + zodd(inZ, b, kz, 4));
    dx = g.fs && kz == 0 ? 0.0f : g.idx * sx;
    dz = g.idz * sz;
}

// D- along x and z at (kx, kz) (src/propagatorStaticFunctions.h:217-231)
__device__ __forceinline__ void dminus2(const Grid &g, const float *inX, const float *inZ, const Cell2 &c, float &dx,
                                        float &dz) {
    const long nz = g.nz, kx = c.kx, kz = c.kz, b = kx * nz;
    if (g.fs && kz == 0) {
        dx = 0.0f;
        dz = 0.0f;
        return;
    }
    const float sx = C1 * (-inX[(kx - 1) * nz + kz] + inX[(kx + 0) * nz + kz]) +
                     C2 * (-inX[(kx - 2) * nz + kz] + inX[(kx + 1) * nz + kz]) +
                     C3 * (-inX[(kx - 3) * nz + kz] + inX[(kx + 2) * nz + kz]) +
                     C4 * (-inX[(kx - 4) * nz + kz] + inX[(kx + 3) * nz + kz]);
    const float sz = C1 * (-zeven(inZ, b, kz, -1) + zeven(inZ, b, kz, 0)) +
                     C2 * (-zeven(inZ, b, kz, -2) + zeven(inZ, b, kz, 1)) +
                     C3 * (-zeven(inZ, b, kz, -3) + zeven(inZ, b, kz, 2)) +
                     C4 * (-zeven(inZ, b, kz, -4) + zeven(inZ, b, kz, 3));
    dx = g.idx * sx;
    dz = g.idz * sz;
}

// generic D+ / D- (src/propagatorStaticFunctions.h:10-311); zero outside the update region
__global__ void __launch_bounds__(256) k2_dplus(Grid g, const float *__restrict__ inX, const float *__restrict__ inZ,
                                                float *__restrict__ outX, float *__restrict__ outZ) {
    const Cell2 c = cell2(g);
    if (!c.in) return;
    float dx = 0.0f, dz = 0.0f;
    if (c.upd) dplus2(g, inX, inZ, c, dx, dz);
    outX[c.k] = dx;
    outZ[c.k] = dz;
}

__global__ void __launch_bounds__(256) k2_dminus(Grid g, const float *__restrict__ inX, const float *__restrict__ inZ,
                                                 float *__restrict__ outX, float *__restrict__ outZ) {
    const Cell2 c = cell2(g);
    if (!c.in) return;
    float dx = 0.0f, dz = 0.0f;
    if (c.upd) dminus2(g, inX, inZ, c, dx, dz);
    outX[c.k] = dx;
    outZ[c.k] = dz;
}

// Iso sandwich: t = b D+ p (src/prop2DAcoIsoDenQ_DEO2_FDTD.h:674-860)
__global__ void __launch_bounds__(256) k2_iso_sandwich(Grid g, const float *__restrict__ p, const float *__restrict__ b,
                                                       float *__restrict__ tX, float *__restrict__ tZ) {
    const Cell2 c = cell2(g);
    if (!c.in) return;
    float ox = 0.0f, oz = 0.0f;
    if (c.upd) {
        float dx, dz;
        dplus2(g, p, p, c, dx, dz);
        const float B = b[c.k];
        ox = B * dx;
        oz = B * dz;
    }
    tX[c.k] = ox;
    tZ[c.k] = oz;
}

// Iso D- and time update (src/prop2DAcoIsoDenQ_DEO2_FDTD.h:863-1045): in the update region
// p+ = dt^2 v^2 / b L - Q (p - p-) - p- + 2 p with L = D-x tX + D-z tZ, pSpace = L; pSpace is 0
// elsewhere and p- is left alone
__global__ void __launch_bounds__(256) k2_iso_update(Grid g, float dt2, const float *__restrict__ tX,
                                                     const float *__restrict__ tZ, const float *__restrict__ v,
                                                     const float *__restrict__ b, const float *__restrict__ q,
                                                     const float *__restrict__ pCur, float *__restrict__ pSpace,
                                                     float *__restrict__ pOld) {
    const Cell2 c = cell2(g);
    if (!c.in) return;
    if (!c.upd) {
        if (pSpace) pSpace[c.k] = 0.0f;
        return;
    }
    float dx, dz;
    dminus2(g, tX, tZ, c, dx, dz);
    const long k = c.k;
    const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
    pOld[k] = dt2V2_B * (dx + dz) - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
    if (pSpace) pSpace[k] = dx + dz;
}

// ---- 2D Iso Born sources and imaging (src/prop2DAcoIsoDenQ_DEO2_FDTD.h:249-582), pointwise
// over the whole grid

// V: p += 2 dt^2 dV / V DP
__global__ void k2_iso_inj_v(long n, float dt, const float *__restrict__ v, const float *__restrict__ dV,
                             const float *__restrict__ DP, float *__restrict__ pCur) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float factor = 2 * dt * dt * dV[k] / v[k];
    pCur[k] += factor * DP[k];
}

// VB (dV given) and B (dV null): p += dt^2 v^2/b (DP/v^2 (2 b dV / v - dB) + lx + lz) or
// p += dt^2 v^2/b (DP (-dB / v^2) + lx + lz), l = D-(dB D+ P)
__global__ void k2_iso_inj_vb(long n, float dt, const float *__restrict__ v, const float *__restrict__ b,
                              const float *__restrict__ dV, const float *__restrict__ dB, const float *__restrict__ DP,
                              const float *__restrict__ lx, const float *__restrict__ lz, float *__restrict__ pCur) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float V = v[k], B = b[k];
    const float V2 = V * V;
    const float dt2v2OverB = dt * dt * V2 / B;
    if (dV) {
        pCur[k] += dt2v2OverB * (DP[k] / V2 * (2 * B * dV[k] / V - dB[k]) + lx[k] + lz[k]);
    } else {
        const float factorB = -dB[k] / V2;
        pCur[k] += dt2v2OverB * (DP[k] * factorB + lx[k] + lz[k]);
    }
}

// t = dB t (the Born buoyancy sandwich's pointwise product)
__global__ void k2_scale2(long n, const float *__restrict__ s, float *__restrict__ a, float *__restrict__ c) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float S = s[k];
    a[k] = S * a[k];
    c[k] = S * c[k];
}

// imaging V / VB / B: gV += 2b/v^3 DP p-; gB += -1/v^2 DP p- - px qx - pz qz (p, q = D+ P, D+ p-)
__global__ void k2_iso_acc(long n, const float *__restrict__ v, const float *__restrict__ b,
                           const float *__restrict__ DP, const float *__restrict__ pOld, const float *__restrict__ px,
                           const float *__restrict__ pz, const float *__restrict__ qx, const float *__restrict__ qz,
                           float *__restrict__ gV, float *__restrict__ gB) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float V = v[k], B = b[k];
    if (gV) {
        const float factorV = +2 * B / (V * V * V);
        gV[k] += factorV * DP[k] * pOld[k];
    }
    if (gB) {
        const float factorB = -1 / (V * V);
        gB[k] += factorB * DP[k] * pOld[k] - px[k] * qx[k] - pz[k] * qz[k];
    }
}

// ---- 2D VTI / TTI (src/prop2DAcoVTIDenQ_DEO2_FDTD.h:686-1477,
// src/prop2DAcoTTIDenQ_DEO2_FDTD.h:667-1717)

// VTI (ROT false) and TTI (ROT true) sandwich: t = material(D+ p, D+ m), the TTI derivatives
// rotated into the symmetry axis' frame first (g1 = cos dx - sin dz, g3 = sin dx + cos dz)
template <bool ROT>
__global__ void __launch_bounds__(256)
    k2_aniso_sandwich(Grid g, const float *__restrict__ p, const float *__restrict__ m, const float *__restrict__ sT,
                      const float *__restrict__ cT, const float *__restrict__ eps, const float *__restrict__ eta,
                      const float *__restrict__ f, const float *__restrict__ b, float *__restrict__ tP1,
                      float *__restrict__ tP3, float *__restrict__ tM1, float *__restrict__ tM3) {
    const Cell2 c = cell2(g);
    if (!c.in) return;
    float o0 = 0.0f, o1 = 0.0f, o2 = 0.0f, o3 = 0.0f;
    if (c.upd) {
        float dpx, dpz, dmx, dmz;
        dplus2(g, p, p, c, dpx, dpz);
        dplus2(g, m, m, c, dmx, dmz);
        const long k = c.k;
        if (ROT) {
            const float s = sT[k], co = cT[k];
            const float p1 = co * dpx - s * dpz, p3 = s * dpx + co * dpz;
            const float m1 = co * dmx - s * dmz, m3 = s * dmx + co * dmz;
            dpx = p1;
            dpz = p3;
            dmx = m1;
            dmz = m3;
        }
        aniso2_material(dpx, dpz, dmx, dmz, eps[k], eta[k], f[k], b[k], o0, o1, o2, o3);
    }
    tP1[c.k] = o0;
    tP3[c.k] = o1;
    tM1[c.k] = o2;
    tM3[c.k] = o3;
}

// weighted backward differences sum_j c_j (-(w a)[-j] + (w a)[j-1]) along x and along z (even
// mirror above a free surface, weight and field both read at the mirrored index)
__device__ __forceinline__ float wminus_x(const float *a, const float *w, long kx, long kz, long nz) {
    const long i0 = (kx - 1) * nz + kz, i1 = kx * nz + kz, i2 = (kx - 2) * nz + kz, i3 = (kx + 1) * nz + kz;
    const long i4 = (kx - 3) * nz + kz, i5 = (kx + 2) * nz + kz, i6 = (kx - 4) * nz + kz, i7 = (kx + 3) * nz + kz;
    return C1 * (-(w[i0] * a[i0]) + w[i1] * a[i1]) + C2 * (-(w[i2] * a[i2]) + w[i3] * a[i3]) +
           C3 * (-(w[i4] * a[i4]) + w[i5] * a[i5]) + C4 * (-(w[i6] * a[i6]) + w[i7] * a[i7]);
}

__device__ __forceinline__ long zeven_idx(long base, long kz, long o) {
    const long zz = kz + o;
    return zz >= 0 ? base + zz : base - zz - 1;
}

__device__ __forceinline__ float wminus_z(const float *a, const float *w, long base, long kz) {
    const long i0 = zeven_idx(base, kz, -1), i1 = zeven_idx(base, kz, 0), i2 = zeven_idx(base, kz, -2),
               i3 = zeven_idx(base, kz, 1), i4 = zeven_idx(base, kz, -3), i5 = zeven_idx(base, kz, 2),
               i6 = zeven_idx(base, kz, -4), i7 = zeven_idx(base, kz, 3);
    return C1 * (-(w[i0] * a[i0]) + w[i1] * a[i1]) + C2 * (-(w[i2] * a[i2]) + w[i3] * a[i3]) +
           C3 * (-(w[i4] * a[i4]) + w[i5] * a[i5]) + C4 * (-(w[i6] * a[i6]) + w[i7] * a[i7]);
}

// TTI D- with the transposed rotation inside the differences (:1531-1716):
// o1 = D-x(cos g1) - D-z(sin g1), o3 = D-x(sin g3) + D-z(cos g3); 0 at the free surface
__device__ __forceinline__ void tti2_minus(const Grid &g, const float *g1, const float *g3, const float *sT,
                                           const float *cT, const Cell2 &c, float &o1, float &o3) {
    if (g.fs && c.kz == 0) {
        o1 = 0.0f;
        o3 = 0.0f;
        return;
    }
    const long nz = g.nz, base = c.kx * nz;
    const float A1 = wminus_x(g1, cT, c.kx, c.kz, nz);
    const float B1 = wminus_z(g1, sT, base, c.kz);
    const float A3 = wminus_x(g3, sT, c.kx, c.kz, nz);
    const float B3 = wminus_z(g3, cT, base, c.kz);
    o1 = g.idx * A1 - g.idz * B1;
    o3 = g.idx * A3 + g.idz * B3;
}

// P/M D- and leapfrog update (VTI :1032-1474, TTI :997-1363); pSpace == null selects the
// linear variant, which neither clears nor writes pSpace/mSpace
template <bool ROT>
__global__ void __launch_bounds__(256)
    k2_aniso_update(Grid g, float dt2, const float *__restrict__ tP1, const float *__restrict__ tP3,
                    const float *__restrict__ tM1, const float *__restrict__ tM3, const float *__restrict__ sT,
                    const float *__restrict__ cT, const float *__restrict__ v, const float *__restrict__ b,
                    const float *__restrict__ q, const float *__restrict__ pCur, const float *__restrict__ mCur,
                    float *__restrict__ pSpace, float *__restrict__ mSpace, float *__restrict__ pOld,
                    float *__restrict__ mOld) {
    const Cell2 c = cell2(g);
    if (!c.in) return;
    const long k = c.k;
    if (!c.upd) {
        if (pSpace) {
            pSpace[k] = 0.0f;
            mSpace[k] = 0.0f;
        }
        return;
    }
    float dP1, dP3, dM1, dM3;
    if (ROT) {
        tti2_minus(g, tP1, tP3, sT, cT, c, dP1, dP3);
        tti2_minus(g, tM1, tM3, sT, cT, c, dM1, dM3);
    } else {
        dminus2(g, tP1, tP3, c, dP1, dP3);
        dminus2(g, tM1, tM3, c, dM1, dM3);
    }
    const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
    const float LP = dP1 + dP3;
    const float LM = dM1 + dM3;
    if (pSpace) {
        pSpace[k] = LP;
        mSpace[k] = LM;
    }
    pOld[k] = dt2V2_B * LP - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
    mOld[k] = dt2V2_B * LM - q[k] * (mCur[k] - mOld[k]) - mOld[k] + 2 * mCur[k];
}

// TTI rotated D+ (:1366-1528); at the free surface g1 = -sin dz, g3 = +cos dz
__global__ void __launch_bounds__(256) k2_tti_dplus(Grid g, const float *__restrict__ g1, const float *__restrict__ g3,
                                                    const float *__restrict__ sT, const float *__restrict__ cT,
                                                    float *__restrict__ o1, float *__restrict__ o3) {
    const Cell2 c = cell2(g);
    if (!c.in) return;
    float a = 0.0f, e = 0.0f;
    if (c.upd) {
        float dx, dz;
        dplus2(g, g1, g3, c, dx, dz);
        const long k = c.k;
        if (g.fs && c.kz == 0) {
            a = -sT[k] * dz;
            e = +cT[k] * dz;
        } else {
            a = cT[k] * dx - sT[k] * dz;
            e = sT[k] * dx + cT[k] * dz;
        }
    }
    o1[c.k] = a;
    o3[c.k] = e;
}

__global__ void __launch_bounds__(256) k2_tti_dminus(Grid g, const float *__restrict__ g1,
                                                     const float *__restrict__ g3, const float *__restrict__ sT,
                                                     const float *__restrict__ cT, float *__restrict__ o1,
                                                     float *__restrict__ o3) {
    const Cell2 c = cell2(g);
    if (!c.in) return;
    float a = 0.0f, e = 0.0f;
    if (c.upd) tti2_minus(g, g1, g3, sT, cT, c, a, e);
    o1[c.k] = a;
    o3[c.k] = e;
}

// 2D VEA source, middle pass (VTI :401-424, TTI :380-403)
__global__ void k2_vea_inj_mid(long n, const float *__restrict__ b, const float *__restrict__ eta,
                               const float *__restrict__ f, const float *__restrict__ dE, const float *__restrict__ dA,
                               const float *__restrict__ tPx1, const float *__restrict__ tPz1,
                               const float *__restrict__ tMz1, float *__restrict__ tPx2, float *__restrict__ tPz2,
                               float *__restrict__ tMx2, float *__restrict__ tMz2) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float A = eta[k], B = b[k], F = f[k], de = dE[k], da = dA[k];
    const float coef = da * B * F * (1 - 2 * A * A) / sqrtf(1 - A * A);
    tPx2[k] = (+2 * B * de) * tPx1[k];
    tPz2[k] = (-2 * B * F * A * da) * tPz1[k] + coef * tMz1[k];
    tMx2[k] = 0;
    tMz2[k] = (+2 * B * F * A * da) * tMz1[k] + coef * tPz1[k];
}

// 2D VEA source, final pass (VTI :438-460, TTI :417-439)
__global__ void k2_vea_inj_final(long n, float dt, const float *__restrict__ v, const float *__restrict__ b,
                                 const float *__restrict__ dV, const float *__restrict__ DP,
                                 const float *__restrict__ DM, const float *__restrict__ tPx,
                                 const float *__restrict__ tPz, const float *__restrict__ tMx,
                                 const float *__restrict__ tMz, float *__restrict__ pCur, float *__restrict__ mCur) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float V = v[k], B = b[k], dv = dV[k];
    const float dt2v2OverB = dt * dt * V * V / B;
    const float factor = 2 * B * dv / (V * V * V);
    pCur[k] += dt2v2OverB * (factor * DP[k] + tPx[k] + tPz[k]);
    mCur[k] += dt2v2OverB * (factor * DM[k] + tMx[k] + tMz[k]);
}

// 2D VEA imaging (VTI :655-684, TTI :631-665): 1 = D+ of the wavefields given, 2 = D+ of
// pOld/mOld; the 2D text scales DP p- and DM m- by the factor separately
__global__ void k2_vea_acc(long n, const float *__restrict__ v, const float *__restrict__ b,
                           const float *__restrict__ eta, const float *__restrict__ f, const float *__restrict__ DP,
                           const float *__restrict__ DM, const float *__restrict__ pOld, const float *__restrict__ mOld,
                           const float *__restrict__ tPx1, const float *__restrict__ tPz1,
                           const float *__restrict__ tMz1, const float *__restrict__ tPx2,
                           const float *__restrict__ tPz2, const float *__restrict__ tMz2, float *__restrict__ gV,
                           float *__restrict__ gE, float *__restrict__ gA) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float V = v[k], A = eta[k], B = b[k], F = f[k];
    const float factor = 2 * B / (V * V * V);
    gV[k] += factor * DP[k] * pOld[k] + factor * DM[k] * mOld[k];
    gE[k] += -2 * B * tPx1[k] * tPx2[k];
    const float cc = B * F * (1 - 2 * A * A) / sqrtf(1 - A * A);
    const float partP = 2 * B * F * A * tPz1[k] - cc * tMz1[k];
    const float partM = 2 * B * F * A * tMz1[k] + cc * tPz1[k];
    gA[k] += partP * tPz2[k] - partM * tMz2[k];
}

// applyFirstDerivatives2D_PlusHalf with caller-given coefficients and spacings
// (src/prop2DAcoVTIDenQ_DEO2_FDTD.cc:154-175, src/propagatorStaticFunctions.h:10-161)
__global__ void __launch_bounds__(256) k2_dplus_coef(Grid g, float4 cc, const float *__restrict__ inX,
                                                     const float *__restrict__ inZ, float *__restrict__ outX,
                                                     float *__restrict__ outZ) {
    const Cell2 c = cell2(g);
    if (!c.in) return;
    float dx = 0.0f, dz = 0.0f;
    if (c.upd) {
        const long nz = g.nz, kx = c.kx, kz = c.kz, b = kx * nz;
        const float sx = cc.x * (-inX[(kx + 0) * nz + kz] + inX[(kx + 1) * nz + kz]) +
                         cc.y * (-inX[(kx - 1) * nz + kz] + inX[(kx + 2) * nz + kz]) +
                         cc.z * (-inX[(kx - 2) * nz + kz] + inX[(kx + 3) * nz + kz]) +
                         cc.w * (-inX[(kx - 3) * nz + kz] + inX[(kx + 4) * nz + kz]);
        const float sz = cc.x * (-zodd(inZ, b, kz, 0) + zodd(inZ, b, kz, 1)) +
                         cc.y * (-zodd(inZ, b, kz, -1) + zodd(inZ, b, kz, 2)) +
                         cc.z * (-zodd(inZ, b, kz, -2) + zodd(inZ, b, kz, 3)) +
                         cc.w * (-zodd(inZ, b, kz, -3) + zodd(inZ, b, kz, 4));
        dx = g.fs && kz == 0 ? 0.0f : g.idx * sx;
        dz = g.idz * sz;
    }
    outX[c.k] = dx;
    outZ[c.k] = dz;
}

// ---- operations ---------------------------------------------------------------------------

// the fused 2D step with a Born epilogue (EPI_SAVE_DP, EPI_BORN_INJ, EPI_IMAGE; arrays as in
// launch_fused2)
template <int PHYS>
static inline bool launch_fused2_epi(Prop *p, int epi, float *e0, float *e1, float *e2) {
    return epi == EPI_SAVE_DP    ? launch_fused2<PHYS, EPI_SAVE_DP>(p, false, e0, e1, e2)
           : epi == EPI_BORN_INJ ? launch_fused2<PHYS, EPI_BORN_INJ>(p, false, e0, e1, e2)
                                 : launch_fused2<PHYS, EPI_IMAGE>(p, false, e0, e1, e2);
}

static inline bool step_2d_epi(Prop *p, int epi, float *e0, float *e1, float *e2) {
    if (p->nx * p->nz >= (1L << 31)) {
        set_error("2D fused step: fewer than 2^31 cells");
        return false;
    }
    return p->phys == ISO   ? launch_fused2_epi<ISO>(p, epi, e0, e1, e2)
           : p->phys == VTI ? launch_fused2_epi<VTI>(p, epi, e0, e1, e2)
                            : launch_fused2_epi<TTI>(p, epi, e0, e1, e2);
}

static inline bool step_2d(Prop *p, bool linear) {
    const Grid &g = p->g;
    Field *f = p->f;
    if (p->fused) {
        if (p->phys == TTI && linear) {
            set_error("2D TTI: there is no linear step");
            return false;
        }
        if (p->nx * p->nz >= (1L << 31)) {
            set_error("2D fused step: fewer than 2^31 cells");
            return false;
        }
        return p->phys == ISO ? launch_fused2<ISO>(p, linear)
                              : (p->phys == VTI ? launch_fused2<VTI>(p, linear) : launch_fused2<TTI>(p, linear));
    }
    if (p->phys == ISO) {
        float *t[2];
        for (int i = 0; i < 2; i++)
            if (!(t[i] = p->scratch(i))) return false;
        k2_iso_sandwich<<<op2_grid(g), op2_block(), 0, p->stream>>>(g, f[F_PCUR].d, f[F_B].d, t[0], t[1]);
        WFD_LAUNCHED();
        k2_iso_update<<<op2_grid(g), op2_block(), 0, p->stream>>>(
            g, p->dt * p->dt, t[0], t[1], f[F_V].d, f[F_B].d, f[F_Q].d, f[F_PCUR].d, linear ? nullptr : f[F_PSPACE].d,
            f[F_POLD].d);
        WFD_LAUNCHED();
        return true;
    }
    if (p->phys == TTI && linear) {
        set_error("2D TTI: there is no linear step");
        return false;
    }
    float *t[4];
    for (int i = 0; i < 4; i++)
        if (!(t[i] = p->scratch(i))) return false;
    const bool rot = p->phys == TTI;
    const float *sT = rot ? f[F_ST].d : nullptr, *cT = rot ? f[F_CT].d : nullptr;
    float *pS = linear ? nullptr : f[F_PSPACE].d, *mS = linear ? nullptr : f[F_MSPACE].d;
    if (rot) {
        k2_aniso_sandwich<true><<<op2_grid(g), op2_block(), 0, p->stream>>>(
            g, f[F_PCUR].d, f[F_MCUR].d, sT, cT, f[F_EPS].d, f[F_ETA].d, f[F_F].d, f[F_B].d, t[0], t[1], t[2], t[3]);
        WFD_LAUNCHED();
        k2_aniso_update<true><<<op2_grid(g), op2_block(), 0, p->stream>>>(
            g, p->dt * p->dt, t[0], t[1], t[2], t[3], sT, cT, f[F_V].d, f[F_B].d, f[F_Q].d, f[F_PCUR].d,
            f[F_MCUR].d, pS, mS, f[F_POLD].d, f[F_MOLD].d);
        WFD_LAUNCHED();
    } else {
        k2_aniso_sandwich<false><<<op2_grid(g), op2_block(), 0, p->stream>>>(
            g, f[F_PCUR].d, f[F_MCUR].d, sT, cT, f[F_EPS].d, f[F_ETA].d, f[F_F].d, f[F_B].d, t[0], t[1], t[2], t[3]);
        WFD_LAUNCHED();
        k2_aniso_update<false><<<op2_grid(g), op2_block(), 0, p->stream>>>(
            g, p->dt * p->dt, t[0], t[1], t[2], t[3], sT, cT, f[F_V].d, f[F_B].d, f[F_Q].d, f[F_PCUR].d,
            f[F_MCUR].d, pS, mS, f[F_POLD].d, f[F_MOLD].d);
        WFD_LAUNCHED();
    }
    return true;
}

// the 2D Q profile (src/absorb.h:23-75): the sponge ramp in double, exp in float, the stored
// value a double expression; computed on the host once and uploaded, the array and the table by
// sponge distance (Prop::qtab, nsponge + 1 values) from the same values
static inline bool setup_q_2d(Prop *p, float freqQ, float qMin, float qInterior) {
    if (!p->begin(bit(F_Q))) return false;
    if ((qMin < FLT_EPSILON) && (qInterior < FLT_EPSILON)) {
        std::printf("Warning -- qMin and qMax unphysical, dtOmegaInvQ set to zero!\n");
        // the reference clears sizeof(float*) bytes, i.e. the first two values (src/absorb.h:44)
        const long nzero = std::min<long>(2, p->n);
        p->q_profile = false;
        WFD_TRY(hipMemsetAsync(p->f[F_Q].d, 0, sizeof(float) * nzero, p->stream));
        return p->end(bit(F_Q), true);
    }
    const long ns = p->nsponge, nx = p->nx, nz = p->nz;
    std::vector<float> qprof(ns);
    // the reference's log(qMin) on a float argument is C++'s float overload (logf), widened to
    // double for the ramp (src/absorb.h:47-48)
    const double lqmin = (double)std::log(qMin), lqmax = (double)std::log(qInterior);
    for (long ks = 0; ks < ns; ks++) {
        const double dk = (double)(ks) / (double)(ns - 1);
        const double lq = lqmin + dk * (lqmax - lqmin);
        qprof[ks] = expf((float)lq);
    }
    const double pi = 3.1415926535897;
    std::vector<float> qt(ns + 1);
    for (long ks = 0; ks < ns; ks++) qt[ks] = (float)(p->dt * 2.0 * pi * freqQ / qprof[ks]);
    qt[ns] = (float)(p->dt * 2.0 * pi * freqQ / qInterior);
    std::vector<float> q((size_t)p->n);
    for (long kx = 0; kx < nx; kx++)
        for (long kz = 0; kz < nz; kz++) {
            const long ksx = std::min(kx, nx - 1 - kx);
            const long ksz = p->fs ? (nz - 1 - kz) : std::min(kz, nz - 1 - kz);
            q[kx * nz + kz] = qt[std::min(std::min(ksx, ksz), ns)];
        }
    if (!p->qtab) WFD_TRY(hipMalloc(&p->qtab, sizeof(float) * (ns + 1)));
    WFD_TRY(hipMemcpyAsync(p->qtab, qt.data(), sizeof(float) * (ns + 1), hipMemcpyHostToDevice, p->stream));
    WFD_TRY(hipMemcpyAsync(p->f[F_Q].d, q.data(), sizeof(float) * p->n, hipMemcpyHostToDevice, p->stream));
    WFD_TRY(hipStreamSynchronize(p->stream));
    p->q_profile = !p->q_exposed;
    return p->end(bit(F_Q), true);
}

// a 2D handle: the 3D handle with ny = 1 (the reference's 2D layout k = kx nz + kz is the 3D
// one with ny = 1), stepped by the 2D kernels
static inline Prop *make_prop_2d(Physics phys, long fs, long nthread, long nx, long nz, long nsponge, float dx,
                                 float dz, float dt, long nbx, long nbz) {
    if (fs > 0 && nz < 8 && nx > 8) {
        // the free-surface rows kz = 0..3 read z-neighbours up to kz + 4 = 7
        set_error("alloc: a free surface needs nz >= 8");
        return nullptr;
    }
    return make_prop(phys, fs, nthread, nx, 1, nz, nsponge, dx, dx, dz, dt, nbx, 1, nbz, 2);
}

static inline bool dplus_2d(Prop *p, const float *in, float *ox, float *oz) {
    k2_dplus<<<op2_grid(p->g), op2_block(), 0, p->stream>>>(p->g, in, in, ox, oz);
    WFD_LAUNCHED();
    return true;
}

static inline bool dminus_2d(Prop *p, const float *ix, const float *iz, float *ox, float *oz) {
    k2_dminus<<<op2_grid(p->g), op2_block(), 0, p->stream>>>(p->g, ix, iz, ox, oz);
    WFD_LAUNCHED();
    return true;
}

// D+ / D- of one field pair, plain (VTI) or rotated (TTI)
static inline bool aniso2_dplus(Prop *p, const float *in, float *o1, float *o3) {
    if (p->phys == TTI) {
        k2_tti_dplus<<<op2_grid(p->g), op2_block(), 0, p->stream>>>(p->g, in, in, p->f[F_ST].d, p->f[F_CT].d, o1, o3);
        WFD_LAUNCHED();
        return true;
    }
    return dplus_2d(p, in, o1, o3);
}

static inline bool aniso2_dminus(Prop *p, const float *i1, const float *i3, float *o1, float *o3) {
    if (p->phys == TTI) {
        k2_tti_dminus<<<op2_grid(p->g), op2_block(), 0, p->stream>>>(p->g, i1, i3, p->f[F_ST].d, p->f[F_CT].d, o1,
                                                                     o3);
        WFD_LAUNCHED();
        return true;
    }
    return dminus_2d(p, i1, i3, o1, o3);
}

}  // namespace wfd
