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

// ---- operations ---------------------------------------------------------------------------

static inline bool step_2d(Prop *p, bool linear) {
    const Grid &g = p->g;
    float *t[2];
    for (int i = 0; i < 2; i++)
        if (!(t[i] = p->scratch(i))) return false;
    Field *f = p->f;
    if (p->phys != ISO) {
        set_error("2D: only the Iso step is built");
        return false;
    }
    k2_iso_sandwich<<<op2_grid(g), op2_block(), 0, p->stream>>>(g, f[F_PCUR].d, f[F_B].d, t[0], t[1]);
    WFD_LAUNCHED();
    k2_iso_update<<<op2_grid(g), op2_block(), 0, p->stream>>>(g, p->dt * p->dt, t[0], t[1], f[F_V].d, f[F_B].d,
                                                              f[F_Q].d, f[F_PCUR].d, linear ? nullptr : f[F_PSPACE].d,
                                                              f[F_POLD].d);
    WFD_LAUNCHED();
    return true;
}

// the 2D Q profile (src/absorb.h:23-75): the sponge ramp in double, exp in float, the stored
// value a double expression; computed on the host once and uploaded
static inline bool setup_q_2d(Prop *p, float freqQ, float qMin, float qInterior) {
    if (!p->begin(bit(F_Q))) return false;
    if ((qMin < FLT_EPSILON) && (qInterior < FLT_EPSILON)) {
        std::printf("Warning -- qMin and qMax unphysical, dtOmegaInvQ set to zero!\n");
        // the reference clears sizeof(float*) bytes, i.e. the first two values (src/absorb.h:44)
        const long nzero = std::min<long>(2, p->n);
        WFD_TRY(hipMemsetAsync(p->f[F_Q].d, 0, sizeof(float) * nzero, p->stream));
        return p->end(bit(F_Q), true);
    }
    const long ns = p->nsponge, nx = p->nx, nz = p->nz;
    std::vector<float> qprof(ns);
    const double lqmin = std::log((double)qMin), lqmax = std::log((double)qInterior);
    for (long ks = 0; ks < ns; ks++) {
        const double dk = (double)(ks) / (double)(ns - 1);
        const double lq = lqmin + dk * (lqmax - lqmin);
        qprof[ks] = expf((float)lq);
    }
    const double pi = 3.1415926535897;
    std::vector<float> q((size_t)p->n);
    for (long kx = 0; kx < nx; kx++)
        for (long kz = 0; kz < nz; kz++) {
            const long ksx = std::min(kx, nx - 1 - kx);
            const long ksz = p->fs ? (nz - 1 - kz) : std::min(kz, nz - 1 - kz);
            const long ks = std::min(ksx, ksz);
            float val = (float)(p->dt * 2.0 * pi * freqQ / qInterior);
            if (ks < ns) val = (float)(p->dt * 2.0 * pi * freqQ / qprof[ks]);
            q[kx * nz + kz] = val;
        }
    WFD_TRY(hipMemcpyAsync(p->f[F_Q].d, q.data(), sizeof(float) * p->n, hipMemcpyHostToDevice, p->stream));
    WFD_TRY(hipStreamSynchronize(p->stream));
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
    Prop *p = make_prop(phys, fs, nthread, nx, 1, nz, nsponge, dx, dx, dz, dt, nbx, 1, nbz);
    if (p) p->dim = 2;
    return p;
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

}  // namespace wfd
