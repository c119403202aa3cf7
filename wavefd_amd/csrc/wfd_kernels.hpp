// wfd_kernels.hpp -- gfx950 device code shared by the WaveFD 3D propagator libraries.
//
// Layout of every field: nx*ny*nz float32, k = kx*ny*nz + ky*nz + kz (z fastest), the
// reference's layout (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:131), so host mirrors and the 1-based
// SourcePoint32 indices of src/spacetime.jl:793-797 map to device memory without a transpose.
//
// This header holds the "operator" kernels: one thread per cell, z along the 64 lanes of a
// wavefront (coalesced), y across the waves of a workgroup, x across workgroups.  Each kernel
// writes its whole output, so the reference's separate annulus-zeroing loops
// (e.g. src/prop3DAcoIsoDenQ_DEO2_FDTD.h:800-838) and free-surface row loops (:895-1031) are
// folded in: a cell is either annulus (output 0), interior [4,n-4)^3, or a free-surface row
// kz=0..3 (only for kx,ky interior).  The fused single-pass time-step kernels live in
// wfd_fused.hpp; these operator kernels serve the Born paths, the two-pass reference step
// used by tests, and any caller of the generic derivative operators.
//
// Floating point: compiled with -ffp-contract=off and IEEE division/sqrt, every expression
// below evaluates the operations, in the order, that the reference source spells out, so the
// results are bit-identical to the CPU restatement of the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wfd {

// 8th-order staggered coefficients, double expressions stored to float
// (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:64-67).
constexpr float C1 = (float)(+1225.0 / 1024.0);
constexpr float C2 = (float)(-245.0 / 3072.0);
constexpr float C3 = (float)(+49.0 / 5120.0);
constexpr float C4 = (float)(-5.0 / 7168.0);

struct Grid {
    long nx, ny, nz, nynz;
    int fs;
    float idx, idy, idz;
};

constexpr int WG_Z = 64;  // lanes along z
constexpr int WG_Y = 4;   // waves along y

inline dim3 op_block() { return dim3(WG_Z, WG_Y, 1); }
inline dim3 op_grid(const Grid &g) {
    return dim3((unsigned)((g.nz + WG_Z - 1) / WG_Z), (unsigned)((g.ny + WG_Y - 1) / WG_Y), (unsigned)g.nx);
}

__device__ __forceinline__ long lmin(long a, long b) { return a < b ? a : b; }

enum Region { R_ANNULUS = 0, R_INTERIOR = 1, R_SURFACE = 2 };

__device__ __forceinline__ int region(const Grid &g, long kx, long ky, long kz) {
    const bool inx = kx >= 4 && kx < g.nx - 4;
    const bool iny = ky >= 4 && ky < g.ny - 4;
    if (inx && iny && kz >= 4 && kz < g.nz - 4) return R_INTERIOR;
    if (g.fs && inx && iny && kz < 4) return R_SURFACE;
    return R_ANNULUS;
}

#define WFD_CELL(g)                                                      \
    const long kz = (long)blockIdx.x * WG_Z + threadIdx.x;               \
    const long ky = (long)blockIdx.y * WG_Y + threadIdx.y;               \
    const long kx = (long)blockIdx.z;                                    \
    if (kz >= (g).nz || ky >= (g).ny) return;                            \
    const long k = kx * (g).nynz + ky * (g).nz + kz;                     \
    const long kc = k - kz;

// forward difference centred at k+1/2 (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:861-874)
__device__ __forceinline__ float sten_p(const float *__restrict__ f, long k, long s) {
    return C1 * (-f[k] + f[k + s]) + C2 * (-f[k - s] + f[k + 2 * s]) +
           C3 * (-f[k - 2 * s] + f[k + 3 * s]) + C4 * (-f[k - 3 * s] + f[k + 4 * s]);
}

// backward difference centred at k-1/2 (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:1127-1144)
__device__ __forceinline__ float sten_m(const float *__restrict__ f, long k, long s) {
    return C1 * (-f[k - s] + f[k]) + C2 * (-f[k - 2 * s] + f[k + s]) +
           C3 * (-f[k - 3 * s] + f[k + 2 * s]) + C4 * (-f[k - 4 * s] + f[k + 3 * s]);
}

// free-surface mirrors: odd f[-j] = -f[j] for D+, even f[-j] = f[j-1] for D-
// (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:905-1029 and :1187-1283)
__device__ __forceinline__ float fodd(const float *__restrict__ col, long j) { return j >= 0 ? col[j] : -col[-j]; }
__device__ __forceinline__ float feven(const float *__restrict__ col, long j) { return j >= 0 ? col[j] : col[-j - 1]; }

__device__ __forceinline__ float sten_p_fs(const float *__restrict__ col, long kz) {
    return C1 * (-fodd(col, kz) + fodd(col, kz + 1)) + C2 * (-fodd(col, kz - 1) + fodd(col, kz + 2)) +
           C3 * (-fodd(col, kz - 2) + fodd(col, kz + 3)) + C4 * (-fodd(col, kz - 3) + fodd(col, kz + 4));
}

__device__ __forceinline__ float sten_m_fs(const float *__restrict__ col, long kz) {
    return C1 * (-feven(col, kz - 1) + feven(col, kz)) + C2 * (-feven(col, kz - 2) + feven(col, kz + 1)) +
           C3 * (-feven(col, kz - 3) + feven(col, kz + 2)) + C4 * (-feven(col, kz - 4) + feven(col, kz + 3));
}

// gradient with the free-surface rules of the D+ sandwiches: at kz=0 the x/y derivatives are 0
__device__ __forceinline__ void grad_plus(const Grid &g, const float *__restrict__ f, long k, long kc, long kz,
                                          int reg, float &dx, float &dy, float &dz) {
    if (reg == R_INTERIOR) {
        dx = g.idx * sten_p(f, k, g.nynz);
        dy = g.idy * sten_p(f, k, g.nz);
        dz = g.idz * sten_p(f, k, 1);
    } else {
        dx = kz == 0 ? 0.0f : g.idx * sten_p(f, k, g.nynz);
        dy = kz == 0 ? 0.0f : g.idy * sten_p(f, k, g.nz);
        dz = g.idz * sten_p_fs(f + kc, kz);
    }
}

// divergence pieces of the D- pass; at the surface row kz=0 all are 0
__device__ __forceinline__ void grad_minus(const Grid &g, const float *__restrict__ fx, const float *__restrict__ fy,
                                           const float *__restrict__ fz, long k, long kc, long kz, int reg,
                                           float &dx, float &dy, float &dz) {
    if (reg == R_INTERIOR) {
        dx = g.idx * sten_m(fx, k, g.nynz);
        dy = g.idy * sten_m(fy, k, g.nz);
        dz = g.idz * sten_m(fz, k, 1);
    } else if (kz == 0) {
        dx = 0.0f;
        dy = 0.0f;
        dz = 0.0f;
    } else {
        dx = g.idx * sten_m(fx, k, g.nynz);
        dy = g.idy * sten_m(fy, k, g.nz);
        dz = g.idz * sten_m_fs(fz + kc, kz);
    }
}

// ---------------------------------------------------------------------------------------
// generic first derivatives (src/propagatorStaticFunctions.h:317-528, 534-738)
// ---------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_dplus(Grid g, const float *__restrict__ inX, const float *__restrict__ inY,
                                               const float *__restrict__ inZ, float *__restrict__ oX,
                                               float *__restrict__ oY, float *__restrict__ oZ) {
    WFD_CELL(g);
    const int reg = region(g, kx, ky, kz);
    if (reg == R_ANNULUS) {
        oX[k] = 0; oY[k] = 0; oZ[k] = 0;
        return;
    }
    if (reg == R_INTERIOR) {
        oX[k] = g.idx * sten_p(inX, k, g.nynz);
        oY[k] = g.idy * sten_p(inY, k, g.nz);
        oZ[k] = g.idz * sten_p(inZ, k, 1);
    } else {
        oX[k] = kz == 0 ? 0.0f : g.idx * sten_p(inX, k, g.nynz);
        oY[k] = kz == 0 ? 0.0f : g.idy * sten_p(inY, k, g.nz);
        oZ[k] = g.idz * sten_p_fs(inZ + kc, kz);
    }
}

__global__ void __launch_bounds__(256) k_dminus(Grid g, const float *__restrict__ inX, const float *__restrict__ inY,
                                                const float *__restrict__ inZ, float *__restrict__ oX,
                                                float *__restrict__ oY, float *__restrict__ oZ) {
    WFD_CELL(g);
    const int reg = region(g, kx, ky, kz);
    if (reg == R_ANNULUS) {
        oX[k] = 0; oY[k] = 0; oZ[k] = 0;
        return;
    }
    float dx, dy, dz;
    grad_minus(g, inX, inY, inZ, k, kc, kz, reg, dx, dy, dz);
    oX[k] = dx; oY[k] = dy; oZ[k] = dz;
}

// ---------------------------------------------------------------------------------------
// material sandwiches
// ---------------------------------------------------------------------------------------

// VTI (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:945-957)
__device__ __forceinline__ void vti_material(float dPx, float dPy, float dPz, float dMx, float dMy, float dMz,
                                             float eps, float eta, float f, float b, float &tPx, float &tPy,
                                             float &tPz, float &tMx, float &tMy, float &tMz) {
    const float E = 1 + 2 * eps;
    const float A = eta;
    const float F = f;
    const float B = b;
    const float SA2 = sqrtf(1 - A * A);
    tPx = B * E * dPx;
    tPy = B * E * dPy;
    tPz = B * (1 - F * A * A) * dPz + B * (F * A * SA2) * dMz;
    tMx = B * (1 - F) * dMx;
    tMy = B * (1 - F) * dMy;
    tMz = B * F * A * SA2 * dPz + B * (1 - F + F * A * A) * dMz;
}

// TTI rotation-sandwich-rotation (src/prop3DAcoTTIDenQ_DEO2_FDTD.h:947-970)
__device__ __forceinline__ void tti_material(float dpdx, float dpdy, float dpdz, float dmdx, float dmdy,
                                             float dmdz, float sT, float cT, float sP, float cP, float eps,
                                             float eta, float f, float b, float &oPx, float &oPy, float &oPz,
                                             float &oMx, float &oMy, float &oMz) {
    const float sTcP = sT * cP;
    const float sTsP = sT * sP;
    const float eta2 = eta * eta;
    const float bF = b * f;
    const float g3P = sTcP * dpdx + sTsP * dpdy + cT * dpdz;
    const float g3M = sTcP * dmdx + sTsP * dmdy + cT * dmdz;
    const float FE = bF * eta * sqrtf(1 - eta2);
    const float tmpP = -b * (2 * eps + f * eta2) * g3P + FE * g3M;
    const float tmpM = FE * g3P + bF * eta2 * g3M;
    const float tmpE = b * (1 + 2 * eps);
    const float tmpF = b * (1 - f);
    oPx = tmpE * dpdx + sTcP * tmpP;
    oPy = tmpE * dpdy + sTsP * tmpP;
    oPz = tmpE * dpdz + cT * tmpP;
    oMx = tmpF * dmdx + sTcP * tmpM;
    oMy = tmpF * dmdy + sTsP * tmpM;
    oMz = tmpF * dmdz + cT * tmpM;
}

// Iso sandwich (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:771-1033)
__global__ void __launch_bounds__(256) k_iso_sandwich(Grid g, const float *__restrict__ p, const float *__restrict__ b,
                                                      float *__restrict__ tX, float *__restrict__ tY,
                                                      float *__restrict__ tZ) {
    WFD_CELL(g);
    const int reg = region(g, kx, ky, kz);
    if (reg == R_ANNULUS) {
        tX[k] = 0; tY[k] = 0; tZ[k] = 0;
        return;
    }
    float dx, dy, dz;
    grad_plus(g, p, k, kc, kz, reg, dx, dy, dz);
    const float B = b[k];
    tX[k] = B * dx;
    tY[k] = B * dy;
    tZ[k] = B * dz;
}

// VTI sandwich (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:793-1212)
__global__ void __launch_bounds__(256) k_vti_sandwich(Grid g, const float *__restrict__ p, const float *__restrict__ m,
                                                      const float *__restrict__ eps, const float *__restrict__ eta,
                                                      const float *__restrict__ f, const float *__restrict__ b,
                                                      float *__restrict__ tPx, float *__restrict__ tPy,
                                                      float *__restrict__ tPz, float *__restrict__ tMx,
                                                      float *__restrict__ tMy, float *__restrict__ tMz) {
    WFD_CELL(g);
    const int reg = region(g, kx, ky, kz);
    if (reg == R_ANNULUS) {
        tPx[k] = 0; tPy[k] = 0; tPz[k] = 0; tMx[k] = 0; tMy[k] = 0; tMz[k] = 0;
        return;
    }
    float dPx, dPy, dPz, dMx, dMy, dMz;
    grad_plus(g, p, k, kc, kz, reg, dPx, dPy, dPz);
    grad_plus(g, m, k, kc, kz, reg, dMx, dMy, dMz);
    vti_material(dPx, dPy, dPz, dMx, dMy, dMz, eps[k], eta[k], f[k], b[k], tPx[k], tPy[k], tPz[k], tMx[k], tMy[k],
                 tMz[k]);
}

// TTI sandwich (src/prop3DAcoTTIDenQ_DEO2_FDTD.h:796-1255)
__global__ void __launch_bounds__(256) k_tti_sandwich(Grid g, const float *__restrict__ p, const float *__restrict__ m,
                                                      const float *__restrict__ sT, const float *__restrict__ cT,
                                                      const float *__restrict__ sP, const float *__restrict__ cP,
                                                      const float *__restrict__ eps, const float *__restrict__ eta,
                                                      const float *__restrict__ f, const float *__restrict__ b,
                                                      float *__restrict__ oPx, float *__restrict__ oPy,
                                                      float *__restrict__ oPz, float *__restrict__ oMx,
                                                      float *__restrict__ oMy, float *__restrict__ oMz) {
    WFD_CELL(g);
    const int reg = region(g, kx, ky, kz);
    if (reg == R_ANNULUS) {
        oPx[k] = 0; oPy[k] = 0; oPz[k] = 0; oMx[k] = 0; oMy[k] = 0; oMz[k] = 0;
        return;
    }
    float dPx, dPy, dPz, dMx, dMy, dMz;
    grad_plus(g, p, k, kc, kz, reg, dPx, dPy, dPz);
    grad_plus(g, m, k, kc, kz, reg, dMx, dMy, dMz);
    tti_material(dPx, dPy, dPz, dMx, dMy, dMz, sT[k], cT[k], sP[k], cP[k], eps[k], eta[k], f[k], b[k], oPx[k],
                 oPy[k], oPz[k], oMx[k], oMy[k], oMz[k]);
}

// ---------------------------------------------------------------------------------------
// D- and second-order time update with Q (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:1039-1284,
// src/prop3DAcoVTIDenQ_DEO2_FDTD.h:1231-1588 nonlinear, :1594-1895 linear)
// ---------------------------------------------------------------------------------------

__device__ __forceinline__ float leapfrog(float dt2V2_B, float L, float q, float cur, float old) {
    return dt2V2_B * L - q * (cur - old) - old + 2 * cur;
}

// pSpace == nullptr: linear variant (no spatial output, no annulus zeroing)
__global__ void __launch_bounds__(256) k_iso_update(Grid g, float dt2, const float *__restrict__ tX,
                                                    const float *__restrict__ tY, const float *__restrict__ tZ,
                                                    const float *__restrict__ v, const float *__restrict__ b,
                                                    const float *__restrict__ q, const float *__restrict__ pCur,
                                                    float *__restrict__ pSpace, float *__restrict__ pOld) {
    WFD_CELL(g);
    const int reg = region(g, kx, ky, kz);
    if (reg == R_ANNULUS) {
        if (pSpace) pSpace[k] = 0;
        return;
    }
    float dx, dy, dz;
    grad_minus(g, tX, tY, tZ, k, kc, kz, reg, dx, dy, dz);
    const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
    const float L = dx + dy + dz;
    pOld[k] = leapfrog(dt2V2_B, L, q[k], pCur[k], pOld[k]);
    if (pSpace) pSpace[k] = L;
}

__global__ void __launch_bounds__(256) k_vti_update(Grid g, float dt2, const float *__restrict__ tPx,
                                                    const float *__restrict__ tPy, const float *__restrict__ tPz,
                                                    const float *__restrict__ tMx, const float *__restrict__ tMy,
                                                    const float *__restrict__ tMz, const float *__restrict__ v,
                                                    const float *__restrict__ b, const float *__restrict__ q,
                                                    const float *__restrict__ pCur, const float *__restrict__ mCur,
                                                    float *__restrict__ pSpace, float *__restrict__ mSpace,
                                                    float *__restrict__ pOld, float *__restrict__ mOld) {
    WFD_CELL(g);
    const int reg = region(g, kx, ky, kz);
    if (reg == R_ANNULUS) {
        if (pSpace) {
            pSpace[k] = 0;
            mSpace[k] = 0;
        }
        return;
    }
    float dPx, dPy, dPz, dMx, dMy, dMz;
    grad_minus(g, tPx, tPy, tPz, k, kc, kz, reg, dPx, dPy, dPz);
    grad_minus(g, tMx, tMy, tMz, k, kc, kz, reg, dMx, dMy, dMz);
    const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
    const float LP = dPx + dPy + dPz;
    const float LM = dMx + dMy + dMz;
    if (pSpace) {
        pSpace[k] = LP;
        mSpace[k] = LM;
    }
    const float qk = q[k];
    pOld[k] = leapfrog(dt2V2_B, LP, qk, pCur[k], pOld[k]);
    mOld[k] = leapfrog(dt2V2_B, LM, qk, mCur[k], mOld[k]);
}

// ---------------------------------------------------------------------------------------
// TTI rotated first derivatives (src/prop3DAcoTTIDenQ_DEO2_FDTD.h:2910-3165, 3171-3657)
// ---------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_tti_dplus_rot(Grid g, const float *__restrict__ g1,
                                                       const float *__restrict__ g2, const float *__restrict__ g3,
                                                       const float *__restrict__ sT, const float *__restrict__ cT,
                                                       const float *__restrict__ sP, const float *__restrict__ cP,
                                                       float *__restrict__ o1, float *__restrict__ o2,
                                                       float *__restrict__ o3) {
    WFD_CELL(g);
    const int reg = region(g, kx, ky, kz);
    if (reg == R_ANNULUS) {
        o1[k] = 0; o2[k] = 0; o3[k] = 0;
        return;
    }
    if (reg == R_SURFACE) {
        const float dz = g.idz * sten_p_fs(g3 + kc, kz);
        if (kz == 0) {
            o1[k] = -sT[k] * dz;
            o2[k] = 0;
            o3[k] = cT[k] * dz;
            return;
        }
        const float dx = g.idx * sten_p(g1, k, g.nynz);
        const float dy = g.idy * sten_p(g2, k, g.nz);
        o1[k] = cT[k] * cP[k] * dx + cT[k] * sP[k] * dy - sT[k] * dz;
        o2[k] = -sP[k] * dx + cP[k] * dy;
        o3[k] = sT[k] * cP[k] * dx + sT[k] * sP[k] * dy + cT[k] * dz;
        return;
    }
    const float dx = g.idx * sten_p(g1, k, g.nynz);
    const float dy = g.idy * sten_p(g2, k, g.nz);
    const float dz = g.idz * sten_p(g3, k, 1);
    const float cTcP = cT[k] * cP[k];
    const float cTsP = cT[k] * sP[k];
    const float sTcP = sT[k] * cP[k];
    o1[k] = cTcP * dx + cTsP * dy - sT[k] * dz;
    o2[k] = -sP[k] * dx + cP[k] * dy;
    o3[k] = sTcP * dx + sT[k] * sP[k] * dy + cT[k] * dz;
}

// weight at neighbour i: (wa*wb)*g or wa*g
__device__ __forceinline__ float wv2(const float *__restrict__ g, const float *__restrict__ wa,
                                     const float *__restrict__ wb, long i) {
    return wa[i] * wb[i] * g[i];
}
__device__ __forceinline__ float wv1(const float *__restrict__ g, const float *__restrict__ wa, long i) {
    return wa[i] * g[i];
}

__device__ __forceinline__ float sten_m_w2(const float *__restrict__ g, const float *__restrict__ wa,
                                           const float *__restrict__ wb, long k, long s) {
    return C1 * (-wv2(g, wa, wb, k - s) + wv2(g, wa, wb, k)) + C2 * (-wv2(g, wa, wb, k - 2 * s) + wv2(g, wa, wb, k + s)) +
           C3 * (-wv2(g, wa, wb, k - 3 * s) + wv2(g, wa, wb, k + 2 * s)) +
           C4 * (-wv2(g, wa, wb, k - 4 * s) + wv2(g, wa, wb, k + 3 * s));
}

__device__ __forceinline__ float sten_m_w1(const float *__restrict__ g, const float *__restrict__ wa, long k, long s) {
    return C1 * (-wv1(g, wa, k - s) + wv1(g, wa, k)) + C2 * (-wv1(g, wa, k - 2 * s) + wv1(g, wa, k + s)) +
           C3 * (-wv1(g, wa, k - 3 * s) + wv1(g, wa, k + 2 * s)) + C4 * (-wv1(g, wa, k - 4 * s) + wv1(g, wa, k + 3 * s));
}

__device__ __forceinline__ long me(long kc, long j) { return kc + (j >= 0 ? j : -j - 1); }

__device__ __forceinline__ float sten_m_w1_fs(const float *__restrict__ g, const float *__restrict__ wa, long kc,
                                              long kz) {
    return C1 * (-wv1(g, wa, me(kc, kz - 1)) + wv1(g, wa, me(kc, kz))) +
           C2 * (-wv1(g, wa, me(kc, kz - 2)) + wv1(g, wa, me(kc, kz + 1))) +
           C3 * (-wv1(g, wa, me(kc, kz - 3)) + wv1(g, wa, me(kc, kz + 2))) +
           C4 * (-wv1(g, wa, me(kc, kz - 4)) + wv1(g, wa, me(kc, kz + 3)));
}

__global__ void __launch_bounds__(256) k_tti_dminus_rot(Grid g, const float *__restrict__ g1,
                                                        const float *__restrict__ g2, const float *__restrict__ g3,
                                                        const float *__restrict__ sT, const float *__restrict__ cT,
                                                        const float *__restrict__ sP, const float *__restrict__ cP,
                                                        float *__restrict__ o1, float *__restrict__ o2,
                                                        float *__restrict__ o3) {
    WFD_CELL(g);
    const int reg = region(g, kx, ky, kz);
    if (reg == R_ANNULUS || (reg == R_SURFACE && kz == 0)) {
        o1[k] = 0; o2[k] = 0; o3[k] = 0;
        return;
    }
    const float A1 = sten_m_w2(g1, cT, cP, k, g.nynz);
    const float B1 = sten_m_w2(g1, cT, sP, k, g.nz);
    const float A2 = sten_m_w1(g2, sP, k, g.nynz);
    const float B2 = sten_m_w1(g2, cP, k, g.nz);
    const float A3 = sten_m_w2(g3, sT, cP, k, g.nynz);
    const float B3 = sten_m_w2(g3, sT, sP, k, g.nz);
    float C1v, C3v;
    if (reg == R_INTERIOR) {
        C1v = sten_m_w1(g1, sT, k, 1);
        C3v = sten_m_w1(g3, cT, k, 1);
    } else {
        C1v = sten_m_w1_fs(g1, sT, kc, kz);
        C3v = sten_m_w1_fs(g3, cT, kc, kz);
    }
    o1[k] = g.idx * A1 + g.idy * B1 - g.idz * C1v;
    o2[k] = -g.idx * A2 + g.idy * B2;
    o3[k] = g.idx * A3 + g.idy * B3 + g.idz * C3v;
}

// ---------------------------------------------------------------------------------------
// sponge / Q profile (src/absorb.h:85-151); qprof is the float profile computed on the
// host exactly as the reference does (logf/expf), the per-cell expression runs in double.
// ---------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_setup_q(Grid g, long nsponge, float dt, float freqQ, float qInterior,
                                                 const float *__restrict__ qprof, float *__restrict__ q) {
    WFD_CELL(g);
    const long ksx = lmin(kx, g.nx - 1 - kx);
    const long ksy = lmin(ky, g.ny - 1 - ky);
    const long ksz = g.fs ? (g.nz - 1 - kz) : lmin(kz, g.nz - 1 - kz);
    const long ks = lmin(ksx, lmin(ksy, ksz));
    const double PI_REF = 3.1415926535897;
    const double num = (double)(dt * 2) * PI_REF * (double)freqQ;
    q[k] = ks < nsponge ? (float)(num / (double)qprof[ks]) : (float)(num / (double)qInterior);
}

}  // namespace wfd
