// wfd_born.hpp -- pointwise Born source / imaging-condition kernels and small utilities.
//
// Each kernel runs over the full grid [0, n) as the reference's loops do (e.g.
// src/prop3DAcoIsoDenQ_DEO2_FDTD.h:305-334 iterates bx from 0 to _nx), and evaluates the
// reference's expressions in the reference's operation order.
#pragma once
#include "wfd_kernels.hpp"

namespace wfd {

#define WFD_PTW(n)                                                   \
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;     \
    if (k >= (n)) return;

constexpr int PW_BLOCK = 256;
inline dim3 pw_grid(long n) { return dim3((unsigned)((n + PW_BLOCK - 1) / PW_BLOCK)); }

struct F3 {
    const float *x, *y, *z;
};
struct W3 {
    float *x, *y, *z;
};

// pSpace *= v^2/b (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:269-291; VTI :384-407 scales mSpace too)
__global__ void k_scale_space(long n, const float *__restrict__ v, const float *__restrict__ b, float *__restrict__ ps,
                              float *__restrict__ ms) {
    WFD_PTW(n);
    const float s = v[k] * v[k] / b[k];
    ps[k] *= s;
    if (ms) ms[k] *= s;
}

// dp = pSpace * v^2/b (dm = mSpace * v^2/b): scaleSpatialDerivatives written to other arrays
__global__ void k_scale_space_into(long n, const float *__restrict__ v, const float *__restrict__ b,
                                   const float *__restrict__ ps, const float *__restrict__ ms, float *__restrict__ dp,
                                   float *__restrict__ dm) {
    WFD_PTW(n);
    const float s = v[k] * v[k] / b[k];
    dp[k] = ps[k] * s;
    if (dm) dm[k] = ms[k] * s;
}

// out = d * in for three components (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:371-373)
__global__ void k_scale3(long n, const float *__restrict__ d, F3 in, W3 out) {
    WFD_PTW(n);
    const float dk = d[k];
    out.x[k] = dk * in.x[k];
    out.y[k] = dk * in.y[k];
    out.z[k] = dk * in.z[k];
}

// Iso V source (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:305-334)
__global__ void k_iso_inj_v(long n, float dt, const float *__restrict__ v, const float *__restrict__ b,
                            const float *__restrict__ dV, const float *__restrict__ DP, float *__restrict__ pCur) {
    WFD_PTW(n);
    const float V = v[k], B = b[k], dv = dV[k];
    const float dt2v2OverB = dt * dt * V * V / B;
    const float factorV = 2 * B * dv / (V * V * V);
    pCur[k] += dt2v2OverB * DP[k] * factorV;
}

// Iso VB / B sources, final pass (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:383-403, 456-474); dV may be null
__global__ void k_iso_inj_vb(long n, float dt, const float *__restrict__ v, const float *__restrict__ b,
                             const float *__restrict__ dV, const float *__restrict__ dB, const float *__restrict__ DP,
                             F3 t, float *__restrict__ pCur) {
    WFD_PTW(n);
    const float V = v[k], B = b[k], db = dB[k];
    const float dt2v2OverB = dt * dt * V * V / B;
    const float factorB = -db / (V * V);
    if (dV) {
        const float factorV = 2 * B * dV[k] / (V * V * V);
        pCur[k] += dt2v2OverB * (DP[k] * (factorV + factorB) + t.x[k] + t.y[k] + t.z[k]);
    } else {
        pCur[k] += dt2v2OverB * (DP[k] * factorB + t.x[k] + t.y[k] + t.z[k]);
    }
}

// Iso V imaging condition (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:518-546)
__global__ void k_iso_acc_v(long n, const float *__restrict__ v, const float *__restrict__ b,
                            const float *__restrict__ DP, const float *__restrict__ pOld, float *__restrict__ gV) {
    WFD_PTW(n);
    const float V = v[k], B = b[k];
    const float factorV = +2 * B / (V * V * V);
    gV[k] += factorV * DP[k] * pOld[k];
}

// Iso VB / B imaging (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:563-656); gV may be null
__global__ void k_iso_acc_vb(long n, const float *__restrict__ v, const float *__restrict__ b,
                             const float *__restrict__ DP, const float *__restrict__ pOld, F3 t1, F3 t2,
                             float *__restrict__ gV, float *__restrict__ gB) {
    WFD_PTW(n);
    const float V = v[k], B = b[k];
    if (gV) {
        const float factorV = +2 * B / (V * V * V);
        gV[k] += factorV * DP[k] * pOld[k];
    }
    const float factorB = -1 / (V * V);
    gB[k] += factorB * DP[k] * pOld[k] - t1.x[k] * t2.x[k] - t1.y[k] * t2.y[k] - t1.z[k] * t2.z[k];
}

// VTI/TTI V source (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:420-452)
__global__ void k_vti_inj_v(long n, float dt, const float *__restrict__ v, const float *__restrict__ dV,
                            const float *__restrict__ DP, const float *__restrict__ DM, float *__restrict__ pCur,
                            float *__restrict__ mCur) {
    WFD_PTW(n);
    const float factor = 2 * dt * dt * dV[k] / v[k];
    pCur[k] += factor * DP[k];
    mCur[k] += factor * DM[k];
}

// VEA source: sandwich of the perturbations (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:493-524)
__global__ void k_vea_inj_mid(long n, const float *__restrict__ b, const float *__restrict__ eta,
                              const float *__restrict__ f, const float *__restrict__ dE, const float *__restrict__ dA,
                              F3 tP1, F3 tM1, W3 tP2, W3 tM2) {
    WFD_PTW(n);
    const float A = eta[k], B = b[k], F = f[k], de = dE[k], da = dA[k];
    const float coef = da * B * F * (1 - 2 * A * A) / sqrtf(1 - A * A);
    tP2.x[k] = (+2 * B * de) * tP1.x[k];
    tP2.y[k] = (+2 * B * de) * tP1.y[k];
    tP2.z[k] = (-2 * B * F * A * da) * tP1.z[k] + coef * tM1.z[k];
    tM2.x[k] = 0;
    tM2.y[k] = 0;
    tM2.z[k] = (+2 * B * F * A * da) * tM1.z[k] + coef * tP1.z[k];
}

// VEA source, final pass (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:535-549)
__global__ void k_vea_inj_final(long n, float dt, const float *__restrict__ v, const float *__restrict__ b,
                                const float *__restrict__ dV, const float *__restrict__ DP,
                                const float *__restrict__ DM, F3 tP, F3 tM, float *__restrict__ pCur,
                                float *__restrict__ mCur) {
    WFD_PTW(n);
    const float V = v[k], B = b[k], dv = dV[k];
    const float dt2v2OverB = dt * dt * V * V / B;
    const float factor = 2 * B * dv / (V * V * V);
    pCur[k] += dt2v2OverB * (factor * DP[k] + tP.x[k] + tP.y[k] + tP.z[k]);
    mCur[k] += dt2v2OverB * (factor * DM[k] + tM.x[k] + tM.y[k] + tM.z[k]);
}

// VTI/TTI V imaging (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:564-589)
__global__ void k_vti_acc_v(long n, const float *__restrict__ v, const float *__restrict__ b,
                            const float *__restrict__ DP, const float *__restrict__ DM, const float *__restrict__ pOld,
                            const float *__restrict__ mOld, float *__restrict__ gV) {
    WFD_PTW(n);
    const float V = v[k], B = b[k];
    const float factor = 2 * B / (V * V * V);
    gV[k] += factor * (DP[k] * pOld[k] + DM[k] * mOld[k]);
}

// VEA imaging (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:756-784)
__global__ void k_vea_acc(long n, const float *__restrict__ v, const float *__restrict__ b,
                          const float *__restrict__ eta, const float *__restrict__ f, const float *__restrict__ DP,
                          const float *__restrict__ DM, const float *__restrict__ pOld, const float *__restrict__ mOld,
                          F3 tP1, F3 tM1, F3 tP2, F3 tM2, float *__restrict__ gV, float *__restrict__ gE,
                          float *__restrict__ gA) {
    WFD_PTW(n);
    const float V = v[k], A = eta[k], B = b[k], F = f[k];
    const float factor = 2 * B / (V * V * V);
    gV[k] += factor * (DP[k] * pOld[k] + DM[k] * mOld[k]);
    gE[k] += (-2 * B * tP1.x[k] * tP2.x[k] - 2 * B * tP1.y[k] * tP2.y[k]);
    const float c = B * F * (1 - 2 * A * A) / sqrtf(1 - A * A);
    const float partP = 2 * B * F * A * tP1.z[k] - c * tM1.z[k];
    const float partM = 2 * B * F * A * tM1.z[k] + c * tP1.z[k];
    gA[k] += (partP * tP2.z[k] - partM * tM2.z[k]);
}

// ---------------------------------------------------------------------------------------
// sparse injection / extraction (src/spacetime.cc:144-204)
//
// Points arrive grouped: all points that share the key (grid index for injection,
// trace index for extraction) are contiguous and in their original order -- the
// reference's blocking guarantees the same (src/spacetime.jl:1087-1129).  One thread owns
// one group and accumulates in order, so there are no atomics and the sum order per key
// equals the reference's.
// ---------------------------------------------------------------------------------------

struct SourcePoint32 {
    long iu;
    long ir;
    float c;
};

// itp (optional): the time index is read from device memory instead of it (graph replays)
__global__ void k_inject_groups(float *__restrict__ field, const float *__restrict__ data, long it, long nt,
                                const SourcePoint32 *__restrict__ pts, const int *__restrict__ gstart, int ngroups,
                                const long *__restrict__ itp) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ngroups) return;
    if (itp) it = *itp;
    const int s = gstart[gi], e = gstart[gi + 1];
    const long iu = pts[s].iu - 1;
    float acc = field[iu];
    for (int i = s; i < e; i++) acc += pts[i].c * data[(pts[i].ir - 1) * nt + it - 1];
    field[iu] = acc;
}

__global__ void k_extract_groups(float *__restrict__ data, const float *__restrict__ field, long it, long nt,
                                 const SourcePoint32 *__restrict__ pts, const int *__restrict__ gstart, int ngroups,
                                 const long *__restrict__ itp) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ngroups) return;
    if (itp) it = *itp;
    const int s = gstart[gi], e = gstart[gi + 1];
    const long id = (pts[s].ir - 1) * nt + it - 1;
    float acc = data[id];
    for (int i = s; i < e; i++) acc += pts[i].c * field[pts[i].iu - 1];
    data[id] = acc;
}

// the device-side time index of a replayed graph
__global__ void k_set_i64(long *p, long v) { *p = v; }
__global__ void k_inc_i64(long *p) { *p += 1; }

// ---- registered points with the time interpolation fused in -----------------------------
// The reference's 8-tap polyphase interpolation (src/spacetime.cc:4-127) maps a coarse trace d
// (nd samples) and a fine one m (nm samples, nh fine per coarse) through filters h[k][0..7].
// Fine sample jk = i*nh + k belongs to row i; the row's taps l in [lo, hi] pair it with coarse
// samples i-3+l, the ranges trimmed at both ends: (0,[3,7]) (1,[2,7]) (2,[1,7]) (3..nd-5,[0,7])
// (nd-4,[0,6]) (nd-3,[0,5]) (nd-2,[0,4]); rows past nd-2 are never visited (nd >= 7 here, where
// every row is visited once).
//   interpadjoint (coarse wavelet -> fine): m[jk] = sum over l = lo..hi of h[k][l] d[i-3+l],
//     accumulated from 0 in l order -- computed for the injected step alone;
//   interpforward (fine traces -> coarse): d[i-3+l] += h[k][l] m[jk] for l = lo..hi, the
//     reference visiting rows, then k, then l; for one coarse sample that is fine-time order, so
//     adding each step's contributions as the step is extracted gives the reference's sums.
__device__ __forceinline__ bool interp_row(long jk, long nh, long nd, long nm, long &i, int &k, int &lo, int &hi) {
    if (jk < 0 || jk >= nm) return false;
    i = jk / nh;
    k = (int)(jk - i * nh);
    if (i > nd - 2) return false;
    lo = i == 0 ? 3 : (i == 1 ? 2 : (i == 2 ? 1 : 0));
    hi = i == nd - 4 ? 6 : (i == nd - 3 ? 5 : (i == nd - 2 ? 4 : 7));
    return true;
}

// injection at fine step it (1-based) of coarse data (nr x nd) through the filters h (nh x 8)
__global__ void k_inject_groups_interp(float *__restrict__ field, const float *__restrict__ data, long it, long nd,
                                       const float *__restrict__ h, long nh, long nm,
                                       const SourcePoint32 *__restrict__ pts, const int *__restrict__ gstart,
                                       int ngroups, const long *__restrict__ itp) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ngroups) return;
    if (itp) it = *itp;
    long i;
    int k, lo, hi;
    if (!interp_row(it - 1, nh, nd, nm, i, k, lo, hi)) return;  // the fine sample is 0: nothing to add
    const int s = gstart[gi], e = gstart[gi + 1];
    const long iu = pts[s].iu - 1;
    float acc = field[iu];
    for (int q = s; q < e; q++) {
        const float *d = data + (pts[q].ir - 1) * nd + (i - 3);
        float m = 0.0f;
        for (int l = lo; l <= hi; l++) m += h[k * 8 + l] * d[l];
        acc += pts[q].c * m;
    }
    field[iu] = acc;
}

// extraction at fine step it into coarse traces (nr x nd): the fine value of the trace, then
// its interpforward contributions
__global__ void k_extract_groups_interp(float *__restrict__ data, const float *__restrict__ field, long it, long nd,
                                        const float *__restrict__ h, long nh, long nm,
                                        const SourcePoint32 *__restrict__ pts, const int *__restrict__ gstart,
                                        int ngroups, const long *__restrict__ itp) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ngroups) return;
    if (itp) it = *itp;
    long i;
    int k, lo, hi;
    if (!interp_row(it - 1, nh, nd, nm, i, k, lo, hi)) return;
    const int s = gstart[gi], e = gstart[gi + 1];
    float v = 0.0f;
    for (int q = s; q < e; q++) v += pts[q].c * field[pts[q].iu - 1];
    float *d = data + (pts[s].ir - 1) * nd + (i - 3);
    for (int l = lo; l <= hi; l++) d[l] += h[k * 8 + l] * v;
}

}  // namespace wfd
