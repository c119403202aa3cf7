/*
 * wavefd_oracle.c -- CPU restatement of WaveFD's 3D acoustic DenQ DEO2 FDTD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is what the HIP product path is compared against,
 * and the "port" CPU baseline that bench.py times.  Nothing in the product
 * (wavefd_amd/, include/) includes, links or calls it.
 *
 * It restates, in plain C99 + OpenMP, the reference C++ kernels listed per function
 * below (paths are relative to the reference tree).  Expression order follows the
 * reference source text so that, built with -ffp-contract=off and without fast-math,
 * every float operation is the one the reference source spells out.
 *
 * Layout: every field is an nx*ny*nz float array indexed k = kx*ny*nz + ky*nz + kz
 * (z fastest), the reference's layout (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:131).
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   wo_dplus / wo_dminus            polynomial exactness of the 8th-order staggered pair,
 *                                   discrete adjointness <D+u,w> = -<u,D-w> incl. the free
 *                                   surface mirrors, constant-field annihilation.
 *   wo_iso_* step                   closed form L(x^2)=2b per axis, energy conservation
 *                                   with Q=0, constant-field invariance of the update.
 *   wo_vti_*                        reduces to the isotropic P update for eps=eta=0,
 *                                   energy conservation, Born linearisation (finite
 *                                   differences in eps and eta).
 *   wo_tti_*                        theta=0 reduces to VTI; theta=90deg reduces to VTI on
 *                                   the transposed grid (pins the rotation convention).
 *   wo_tti_dplus_rot/dminus_rot     adjointness <R+u,w> = -<u,R-w>; reduce to wo_dplus /
 *                                   wo_dminus at theta=0.
 *   wo_setup_q                      closed form of the log-linear Q ramp; the reference's
 *                                   own absorb tests (test/absorb.jl:40-99).
 *   wo_inject / wo_extract / interp bit-exact against the reference spacetime.cc compiled
 *                                   into oracle/_ref (oracle/Makefile), plus the reference
 *                                   dot-product test (test/spacetime.jl:390-414).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <float.h>
#include <stddef.h>
#include <omp.h>
#include <xmmintrin.h>

/* Stencil coefficients: double expressions stored to float, as the reference's
 * constructor initialiser list does (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:64-67). */
#define WC1 ((float)(+1225.0 / 1024.0))
#define WC2 ((float)(-245.0 / 3072.0))
#define WC3 ((float)(+49.0 / 5120.0))
#define WC4 ((float)(-5.0 / 7168.0))

/* PI as the reference's absorb.h spells it (src/absorb.h:12). */
#define WPI 3.1415926535897

static inline long imin(long a, long b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------- */
/* stencil building blocks                                                    */
/* ------------------------------------------------------------------------- */

/* forward-staggered 8th-order difference centred at k+1/2 along stride s
 * (e.g. src/prop3DAcoIsoDenQ_DEO2_FDTD.h:861-865) */
static inline float sten_plus(const float *f, long k, long s) {
    return WC1 * (-f[k] + f[k + s]) +
           WC2 * (-f[k - s] + f[k + 2 * s]) +
           WC3 * (-f[k - 2 * s] + f[k + 3 * s]) +
           WC4 * (-f[k - 3 * s] + f[k + 4 * s]);
}

/* backward-staggered 8th-order difference centred at k-1/2 along stride s
 * (e.g. src/prop3DAcoIsoDenQ_DEO2_FDTD.h:1127-1131) */
static inline float sten_minus(const float *f, long k, long s) {
    return WC1 * (-f[k - s] + f[k]) +
           WC2 * (-f[k - 2 * s] + f[k + s]) +
           WC3 * (-f[k - 3 * s] + f[k + 2 * s]) +
           WC4 * (-f[k - 4 * s] + f[k + 3 * s]);
}

/* Odd mirror above the free surface for the forward difference: f[-j] = -f[j].
 * This reproduces the hand-unrolled rows kz=0..3 of the reference
 * (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:905-1029): e.g. at kz=0 the c8_2 term reads
 * (+f1 + f2), which is -(-f1) + f2. */
static inline float fs_odd(const float *col, long j) { return j >= 0 ? col[j] : -col[-j]; }

/* Even mirror for the backward difference: f[-j] = f[j-1]
 * (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:1187-1283). */
static inline float fs_even(const float *col, long j) { return j >= 0 ? col[j] : col[-j - 1]; }

static inline float sten_plus_fs(const float *col, long kz) {
    return WC1 * (-fs_odd(col, kz) + fs_odd(col, kz + 1)) +
           WC2 * (-fs_odd(col, kz - 1) + fs_odd(col, kz + 2)) +
           WC3 * (-fs_odd(col, kz - 2) + fs_odd(col, kz + 3)) +
           WC4 * (-fs_odd(col, kz - 3) + fs_odd(col, kz + 4));
}

static inline float sten_minus_fs(const float *col, long kz) {
    return WC1 * (-fs_even(col, kz - 1) + fs_even(col, kz)) +
           WC2 * (-fs_even(col, kz - 2) + fs_even(col, kz + 1)) +
           WC3 * (-fs_even(col, kz - 3) + fs_even(col, kz + 2)) +
           WC4 * (-fs_even(col, kz - 4) + fs_even(col, kz + 3));
}

/* Zero the 4-cell annulus on all six faces (e.g. src/prop3DAcoIsoDenQ_DEO2_FDTD.h:800-838). */
static void zero_annulus(long nx, long ny, long nz, float **arrs, int narr) {
    const long nynz = ny * nz;
#pragma omp parallel for schedule(static)
    for (long kx = 0; kx < nx; kx++) {
        if (kx < 4 || kx >= nx - 4) {
            for (int a = 0; a < narr; a++) memset(arrs[a] + kx * nynz, 0, sizeof(float) * nynz);
            continue;
        }
        for (long ky = 0; ky < ny; ky++) {
            float *row;
            if (ky < 4 || ky >= ny - 4) {
                for (int a = 0; a < narr; a++) memset(arrs[a] + kx * nynz + ky * nz, 0, sizeof(float) * nz);
                continue;
            }
            for (int a = 0; a < narr; a++) {
                row = arrs[a] + kx * nynz + ky * nz;
                for (long kz = 0; kz < nz; kz++)
                    if (kz < 4 || kz >= nz - 4) row[kz] = 0;
            }
        }
    }
}

/* Iterate the interior box [4,n-4)^3 in the reference's cache blocks (nbx,nby,nbz);
 * blocking never changes results, only the traversal. */
#define FOR_INTERIOR(nx, ny, nz, nbx, nby, nbz, nthread, ...)                                 \
    do {                                                                                        \
        const long _nx4 = (nx) - 4, _ny4 = (ny) - 4, _nz4 = (nz) - 4;                           \
        _Pragma("omp parallel for collapse(3) num_threads(nthread) schedule(static)")           \
        for (long bx = 4; bx < _nx4; bx += (nbx))                                               \
            for (long by = 4; by < _ny4; by += (nby))                                           \
                for (long bz = 4; bz < _nz4; bz += (nbz)) {                                     \
                    const long kxmax = imin(bx + (nbx), _nx4);                                  \
                    const long kymax = imin(by + (nby), _ny4);                                  \
                    const long kzmax = imin(bz + (nbz), _nz4);                                  \
                    for (long kx = bx; kx < kxmax; kx++)                                        \
                        for (long ky = by; ky < kymax; ky++) {                                  \
                            _Pragma("omp simd")                                                 \
                            for (long kz = bz; kz < kzmax; kz++) {                              \
                                const long k = kx * nynz + ky * nz + kz;                        \
                                __VA_ARGS__                                                     \
                            }                                                                   \
                        }                                                                       \
                }                                                                               \
    } while (0)

/* Iterate the free-surface rows kz=0..3 over kx in [4,nx-4), ky in [4,ny-4). */
#define FOR_FS_ROWS(nx, ny, nz, nthread, ...)                                                  \
    do {                                                                                        \
        _Pragma("omp parallel for num_threads(nthread) schedule(static)")                       \
        for (long kx = 4; kx < (nx) - 4; kx++)                                                  \
            for (long ky = 4; ky < (ny) - 4; ky++)                                              \
                for (long kz = 0; kz < 4; kz++) {                                               \
                    const long k = kx * nynz + ky * nz + kz;                                    \
                    const long kc = kx * nynz + ky * nz;                                        \
                    __VA_ARGS__                                                                 \
                }                                                                               \
    } while (0)

/* ------------------------------------------------------------------------- */
/* Q profile: src/absorb.h:85-151                                             */
/* ------------------------------------------------------------------------- */

/* Returns 0, or 1 when freqQ < FLT_EPSILON (where the reference exits the process,
 * src/absorb.h:98-103). */
int wo_setup_q(long fs, long nx, long ny, long nz, long nsponge, float dt,
               float freqQ, float qMin, float qInterior, float *q) {
    if (freqQ < FLT_EPSILON) return 1;
    if ((qMin < FLT_EPSILON) && (qInterior < FLT_EPSILON)) {
        /* the reference memsets sizeof(pointer) bytes, i.e. the first two floats
         * (src/absorb.h:107) */
        long nzero = imin(2, nx * ny * nz);
        for (long i = 0; i < nzero; i++) q[i] = 0;
        return 0;
    }
    const long nynz = ny * nz;
    float *qprof = (float *)malloc(sizeof(float) * (nsponge > 0 ? nsponge : 1));
    const float lqmin = logf(qMin);
    const float lqmax = logf(qInterior);
    for (long ks = 0; ks < nsponge; ks++) {
        const float dk = (float)(ks) / (float)(nsponge - 1);
        const float lq = lqmin + dk * (lqmax - lqmin);
        qprof[ks] = expf(lq);
    }
    for (long kz = 0; kz < nz; kz++)
        for (long kx = 0; kx < nx; kx++)
            for (long ky = 0; ky < ny; ky++) {
                const long ksx = imin(kx, nx - 1 - kx);
                const long ksy = imin(ky, ny - 1 - ky);
                const long ksz = fs ? (nz - 1 - kz) : imin(kz, nz - 1 - kz);
                const long ks = imin(ksx, imin(ksy, ksz));
                const long k = kx * nynz + ky * nz + kz;
                /* dt * 2 is float, then the double PI promotes the rest (src/absorb.h:141-144) */
                q[k] = (float)((double)(dt * 2) * WPI * (double)freqQ / (double)qInterior);
                if (ks < nsponge) q[k] = (float)((double)(dt * 2) * WPI * (double)freqQ / (double)qprof[ks]);
            }
    free(qprof);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Generic first derivatives: src/propagatorStaticFunctions.h:317-528, 534-738 */
/* ------------------------------------------------------------------------- */

void wo_dplus(long fs, long nx, long ny, long nz, float idx, float idy, float idz,
              const float *inX, const float *inY, const float *inZ,
              float *outX, float *outY, float *outZ, long nbx, long nby, long nbz, long nthread) {
    const long nynz = ny * nz;
    float *outs[3] = {outX, outY, outZ};
    zero_annulus(nx, ny, nz, outs, 3);
    FOR_INTERIOR(nx, ny, nz, nbx, nby, nbz, nthread, {
        outX[k] = idx * sten_plus(inX, k, nynz);
        outY[k] = idy * sten_plus(inY, k, nz);
        outZ[k] = idz * sten_plus(inZ, k, 1);
    });
    if (fs) {
        FOR_FS_ROWS(nx, ny, nz, nthread, {
            if (kz == 0) {
                outX[k] = 0;
                outY[k] = 0;
            } else {
                outX[k] = idx * sten_plus(inX, k, nynz);
                outY[k] = idy * sten_plus(inY, k, nz);
            }
            outZ[k] = idz * sten_plus_fs(inZ + kc, kz);
        });
    }
}

void wo_dminus(long fs, long nx, long ny, long nz, float idx, float idy, float idz,
               const float *inX, const float *inY, const float *inZ,
               float *outX, float *outY, float *outZ, long nbx, long nby, long nbz, long nthread) {
    const long nynz = ny * nz;
    float *outs[3] = {outX, outY, outZ};
    zero_annulus(nx, ny, nz, outs, 3);
    FOR_INTERIOR(nx, ny, nz, nbx, nby, nbz, nthread, {
        outX[k] = idx * sten_minus(inX, k, nynz);
        outY[k] = idy * sten_minus(inY, k, nz);
        outZ[k] = idz * sten_minus(inZ, k, 1);
    });
    if (fs) {
        FOR_FS_ROWS(nx, ny, nz, nthread, {
            if (kz == 0) {
                outX[k] = 0;
                outY[k] = 0;
                outZ[k] = 0;
            } else {
                outX[k] = idx * sten_minus(inX, k, nynz);
                outY[k] = idy * sten_minus(inY, k, nz);
                outZ[k] = idz * sten_minus_fs(inZ + kc, kz);
            }
        });
    }
}

/* ------------------------------------------------------------------------- */
/* Isotropic: src/prop3DAcoIsoDenQ_DEO2_FDTD.h:771-1033 and 1039-1284          */
/* ------------------------------------------------------------------------- */

void wo_iso_sandwich(long fs, long nx, long ny, long nz, float idx, float idy, float idz,
                     const float *p, const float *b, float *tX, float *tY, float *tZ,
                     long nbx, long nby, long nbz, long nthread) {
    const long nynz = ny * nz;
    float *outs[3] = {tX, tY, tZ};
    zero_annulus(nx, ny, nz, outs, 3);
    FOR_INTERIOR(nx, ny, nz, nbx, nby, nbz, nthread, {
        const float dPx = idx * sten_plus(p, k, nynz);
        const float dPy = idy * sten_plus(p, k, nz);
        const float dPz = idz * sten_plus(p, k, 1);
        const float B = b[k];
        tX[k] = B * dPx;
        tY[k] = B * dPy;
        tZ[k] = B * dPz;
    });
    if (fs) {
        FOR_FS_ROWS(nx, ny, nz, nthread, {
            const float dPx = kz == 0 ? 0.0f : idx * sten_plus(p, k, nynz);
            const float dPy = kz == 0 ? 0.0f : idy * sten_plus(p, k, nz);
            const float dPz = idz * sten_plus_fs(p + kc, kz);
            const float B = b[k];
            tX[k] = B * dPx;
            tY[k] = B * dPy;
            tZ[k] = B * dPz;
        });
    }
}

/* pSpace may be NULL (linear variant: no spatial output, no annulus zeroing,
 * src/prop3DAcoVTIDenQ_DEO2_FDTD.h:1594-1895). */
void wo_iso_update(long fs, long nx, long ny, long nz, float idx, float idy, float idz, float dt,
                   const float *tX, const float *tY, const float *tZ,
                   const float *v, const float *b, const float *q, const float *pCur,
                   float *pSpace, float *pOld, long nbx, long nby, long nbz, long nthread) {
    const long nynz = ny * nz;
    const float dt2 = dt * dt;
    if (pSpace) {
        float *outs[1] = {pSpace};
        zero_annulus(nx, ny, nz, outs, 1);
    }
    FOR_INTERIOR(nx, ny, nz, nbx, nby, nbz, nthread, {
        const float dPx = idx * sten_minus(tX, k, nynz);
        const float dPy = idy * sten_minus(tY, k, nz);
        const float dPz = idz * sten_minus(tZ, k, 1);
        const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
        pOld[k] = dt2V2_B * (dPx + dPy + dPz) - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
        if (pSpace) pSpace[k] = dPx + dPy + dPz;
    });
    if (fs) {
        FOR_FS_ROWS(nx, ny, nz, nthread, {
            float dPx = 0, dPy = 0, dPz = 0;
            if (kz > 0) {
                dPx = idx * sten_minus(tX, k, nynz);
                dPy = idy * sten_minus(tY, k, nz);
                dPz = idz * sten_minus_fs(tZ + kc, kz);
            }
            const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
            pOld[k] = dt2V2_B * (dPx + dPy + dPz) - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
            if (pSpace) pSpace[k] = dPx + dPy + dPz;
        });
    }
}

/* ------------------------------------------------------------------------- */
/* VTI: src/prop3DAcoVTIDenQ_DEO2_FDTD.h:793-1212, 1231-1588, 1594-1895        */
/* ------------------------------------------------------------------------- */

static inline void vti_material(float dPx, float dPy, float dPz, float dMx, float dMy, float dMz,
                                float eps, float eta, float f, float b,
                                float *tPx, float *tPy, float *tPz, float *tMx, float *tMy, float *tMz) {
    /* src/prop3DAcoVTIDenQ_DEO2_FDTD.h:945-957 */
    const float E = 1 + 2 * eps;
    const float A = eta;
    const float F = f;
    const float B = b;
    const float SA2 = sqrtf(1 - A * A);
    *tPx = B * E * dPx;
    *tPy = B * E * dPy;
    *tPz = B * (1 - F * A * A) * dPz + B * (F * A * SA2) * dMz;
    *tMx = B * (1 - F) * dMx;
    *tMy = B * (1 - F) * dMy;
    *tMz = B * F * A * SA2 * dPz + B * (1 - F + F * A * A) * dMz;
}

void wo_vti_sandwich(long fs, long nx, long ny, long nz, float idx, float idy, float idz,
                     const float *p, const float *m, const float *eps, const float *eta,
                     const float *f, const float *b,
                     float *tPx, float *tPy, float *tPz, float *tMx, float *tMy, float *tMz,
                     long nbx, long nby, long nbz, long nthread) {
    const long nynz = ny * nz;
    float *outs[6] = {tPx, tPy, tPz, tMx, tMy, tMz};
    zero_annulus(nx, ny, nz, outs, 6);
    FOR_INTERIOR(nx, ny, nz, nbx, nby, nbz, nthread, {
        const float dPx = idx * sten_plus(p, k, nynz);
        const float dPy = idy * sten_plus(p, k, nz);
        const float dPz = idz * sten_plus(p, k, 1);
        const float dMx = idx * sten_plus(m, k, nynz);
        const float dMy = idy * sten_plus(m, k, nz);
        const float dMz = idz * sten_plus(m, k, 1);
        vti_material(dPx, dPy, dPz, dMx, dMy, dMz, eps[k], eta[k], f[k], b[k],
                     &tPx[k], &tPy[k], &tPz[k], &tMx[k], &tMy[k], &tMz[k]);
    });
    if (fs) {
        FOR_FS_ROWS(nx, ny, nz, nthread, {
            const float dPx = kz == 0 ? 0.0f : idx * sten_plus(p, k, nynz);
            const float dPy = kz == 0 ? 0.0f : idy * sten_plus(p, k, nz);
            const float dPz = idz * sten_plus_fs(p + kc, kz);
            const float dMx = kz == 0 ? 0.0f : idx * sten_plus(m, k, nynz);
            const float dMy = kz == 0 ? 0.0f : idy * sten_plus(m, k, nz);
            const float dMz = idz * sten_plus_fs(m + kc, kz);
            vti_material(dPx, dPy, dPz, dMx, dMy, dMz, eps[k], eta[k], f[k], b[k],
                         &tPx[k], &tPy[k], &tPz[k], &tMx[k], &tMy[k], &tMz[k]);
        });
    }
}

/* Coupled P/M D- and leapfrog update.  pSpace/mSpace NULL selects the linear variant.
 * The TTI nonlinear update is the same text (src/prop3DAcoTTIDenQ_DEO2_FDTD.h:1274-1631). */
void wo_vti_update(long fs, long nx, long ny, long nz, float idx, float idy, float idz, float dt,
                   const float *tPx, const float *tPy, const float *tPz,
                   const float *tMx, const float *tMy, const float *tMz,
                   const float *v, const float *b, const float *q,
                   const float *pCur, const float *mCur, float *pSpace, float *mSpace,
                   float *pOld, float *mOld, long nbx, long nby, long nbz, long nthread) {
    const long nynz = ny * nz;
    const float dt2 = dt * dt;
    if (pSpace) {
        float *outs[2] = {pSpace, mSpace};
        zero_annulus(nx, ny, nz, outs, 2);
    }
    FOR_INTERIOR(nx, ny, nz, nbx, nby, nbz, nthread, {
        const float dPx = idx * sten_minus(tPx, k, nynz);
        const float dPy = idy * sten_minus(tPy, k, nz);
        const float dPz = idz * sten_minus(tPz, k, 1);
        const float dMx = idx * sten_minus(tMx, k, nynz);
        const float dMy = idy * sten_minus(tMy, k, nz);
        const float dMz = idz * sten_minus(tMz, k, 1);
        const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
        const float LP = dPx + dPy + dPz;
        const float LM = dMx + dMy + dMz;
        if (pSpace) { pSpace[k] = LP; mSpace[k] = LM; }
        pOld[k] = dt2V2_B * LP - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
        mOld[k] = dt2V2_B * LM - q[k] * (mCur[k] - mOld[k]) - mOld[k] + 2 * mCur[k];
    });
    if (fs) {
        FOR_FS_ROWS(nx, ny, nz, nthread, {
            float dPx = 0, dPy = 0, dPz = 0, dMx = 0, dMy = 0, dMz = 0;
            if (kz > 0) {
                dPx = idx * sten_minus(tPx, k, nynz);
                dPy = idy * sten_minus(tPy, k, nz);
                dPz = idz * sten_minus_fs(tPz + kc, kz);
                dMx = idx * sten_minus(tMx, k, nynz);
                dMy = idy * sten_minus(tMy, k, nz);
                dMz = idz * sten_minus_fs(tMz + kc, kz);
            }
            const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
            const float LP = dPx + dPy + dPz;
            const float LM = dMx + dMy + dMz;
            if (pSpace) { pSpace[k] = LP; mSpace[k] = LM; }
            pOld[k] = dt2V2_B * LP - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
            mOld[k] = dt2V2_B * LM - q[k] * (mCur[k] - mOld[k]) - mOld[k] + 2 * mCur[k];
        });
    }
}

/* ------------------------------------------------------------------------- */
/* TTI: src/prop3DAcoTTIDenQ_DEO2_FDTD.h:796-1255 (rotation sandwich)          */
/* ------------------------------------------------------------------------- */

static inline void tti_material(float dpdx, float dpdy, float dpdz, float dmdx, float dmdy, float dmdz,
                                float sT, float cT, float sP, float cP,
                                float eps, float eta, float f, float b,
                                float *oPx, float *oPy, float *oPz, float *oMx, float *oMy, float *oMz) {
    /* src/prop3DAcoTTIDenQ_DEO2_FDTD.h:947-970 */
    const float sinThetaCosPhi = sT * cP;
    const float sinThetaSinPhi = sT * sP;
    const float eta2 = eta * eta;
    const float bF = b * f;
    const float g3P = sinThetaCosPhi * dpdx + sinThetaSinPhi * dpdy + cT * dpdz;
    const float g3M = sinThetaCosPhi * dmdx + sinThetaSinPhi * dmdy + cT * dmdz;
    const float FE = bF * eta * sqrtf(1 - eta2);
    const float tmpP = -b * (2 * eps + f * eta2) * g3P + FE * g3M;
    const float tmpM = FE * g3P + bF * eta2 * g3M;
    const float tmpE = b * (1 + 2 * eps);
    const float tmpF = b * (1 - f);
    *oPx = tmpE * dpdx + sinThetaCosPhi * tmpP;
    *oPy = tmpE * dpdy + sinThetaSinPhi * tmpP;
    *oPz = tmpE * dpdz + cT * tmpP;
    *oMx = tmpF * dmdx + sinThetaCosPhi * tmpM;
    *oMy = tmpF * dmdy + sinThetaSinPhi * tmpM;
    *oMz = tmpF * dmdz + cT * tmpM;
}

void wo_tti_sandwich(long fs, long nx, long ny, long nz, float idx, float idy, float idz,
                     const float *p, const float *m,
                     const float *sT, const float *cT, const float *sP, const float *cP,
                     const float *eps, const float *eta, const float *f, const float *b,
                     float *oPx, float *oPy, float *oPz, float *oMx, float *oMy, float *oMz,
                     long nbx, long nby, long nbz, long nthread) {
    const long nynz = ny * nz;
    float *outs[6] = {oPx, oPy, oPz, oMx, oMy, oMz};
    zero_annulus(nx, ny, nz, outs, 6);
    FOR_INTERIOR(nx, ny, nz, nbx, nby, nbz, nthread, {
        const float dpdx = idx * sten_plus(p, k, nynz);
        const float dpdy = idy * sten_plus(p, k, nz);
        const float dpdz = idz * sten_plus(p, k, 1);
        const float dmdx = idx * sten_plus(m, k, nynz);
        const float dmdy = idy * sten_plus(m, k, nz);
        const float dmdz = idz * sten_plus(m, k, 1);
        tti_material(dpdx, dpdy, dpdz, dmdx, dmdy, dmdz, sT[k], cT[k], sP[k], cP[k],
                     eps[k], eta[k], f[k], b[k], &oPx[k], &oPy[k], &oPz[k], &oMx[k], &oMy[k], &oMz[k]);
    });
    if (fs) {
        FOR_FS_ROWS(nx, ny, nz, nthread, {
            const float dpdx = kz == 0 ? 0.0f : idx * sten_plus(p, k, nynz);
            const float dpdy = kz == 0 ? 0.0f : idy * sten_plus(p, k, nz);
            const float dpdz = idz * sten_plus_fs(p + kc, kz);
            const float dmdx = kz == 0 ? 0.0f : idx * sten_plus(m, k, nynz);
            const float dmdy = kz == 0 ? 0.0f : idy * sten_plus(m, k, nz);
            const float dmdz = idz * sten_plus_fs(m + kc, kz);
            tti_material(dpdx, dpdy, dpdz, dmdx, dmdy, dmdz, sT[k], cT[k], sP[k], cP[k],
                         eps[k], eta[k], f[k], b[k], &oPx[k], &oPy[k], &oPz[k], &oMx[k], &oMy[k], &oMz[k]);
        });
    }
}

/* Rotated D+ : src/prop3DAcoTTIDenQ_DEO2_FDTD.h:2910-3165 */
void wo_tti_dplus_rot(long fs, long nx, long ny, long nz, float idx, float idy, float idz,
                      const float *g1, const float *g2, const float *g3,
                      const float *sT, const float *cT, const float *sP, const float *cP,
                      float *o1, float *o2, float *o3, long nbx, long nby, long nbz, long nthread) {
    const long nynz = ny * nz;
    float *outs[3] = {o1, o2, o3};
    zero_annulus(nx, ny, nz, outs, 3);
    FOR_INTERIOR(nx, ny, nz, nbx, nby, nbz, nthread, {
        const float dx = idx * sten_plus(g1, k, nynz);
        const float dy = idy * sten_plus(g2, k, nz);
        const float dz = idz * sten_plus(g3, k, 1);
        const float cTcP = cT[k] * cP[k];
        const float cTsP = cT[k] * sP[k];
        const float sTcP = sT[k] * cP[k];
        o1[k] = cTcP * dx + cTsP * dy - sT[k] * dz;
        o2[k] = -sP[k] * dx + cP[k] * dy;
        o3[k] = sTcP * dx + sT[k] * sP[k] * dy + cT[k] * dz;
    });
    if (fs) {
        FOR_FS_ROWS(nx, ny, nz, nthread, {
            const float dz = idz * sten_plus_fs(g3 + kc, kz);
            if (kz == 0) {
                o1[k] = -sT[k] * dz;
                o2[k] = 0;
                o3[k] = cT[k] * dz;
            } else {
                const float dx = idx * sten_plus(g1, k, nynz);
                const float dy = idy * sten_plus(g2, k, nz);
                o1[k] = cT[k] * cP[k] * dx + cT[k] * sP[k] * dy - sT[k] * dz;
                o2[k] = -sP[k] * dx + cP[k] * dy;
                o3[k] = sT[k] * cP[k] * dx + sT[k] * sP[k] * dy + cT[k] * dz;
            }
        });
    }
}

/* weighted backward difference: sum_j c_j (-(w g)[k-j] + (w g)[k+j-1]) where the weight
 * w = wa*wb (wb may be NULL) multiplies the field at each neighbour
 * (src/prop3DAcoTTIDenQ_DEO2_FDTD.h:3272-3323). zidx maps a neighbour offset to the
 * index used (even mirror at the free surface). */
static inline float wval(const float *g, const float *wa, const float *wb, long i) {
    return wb ? wa[i] * wb[i] * g[i] : wa[i] * g[i];
}

static inline float sten_minus_w(const float *g, const float *wa, const float *wb, long k, long s) {
    return WC1 * (-wval(g, wa, wb, k - s) + wval(g, wa, wb, k)) +
           WC2 * (-wval(g, wa, wb, k - 2 * s) + wval(g, wa, wb, k + s)) +
           WC3 * (-wval(g, wa, wb, k - 3 * s) + wval(g, wa, wb, k + 2 * s)) +
           WC4 * (-wval(g, wa, wb, k - 4 * s) + wval(g, wa, wb, k + 3 * s));
}

static inline long mirror_even(long kc, long j) { return kc + (j >= 0 ? j : -j - 1); }

static inline float sten_minus_w_fs(const float *g, const float *wa, long kc, long kz) {
    return WC1 * (-wval(g, wa, NULL, mirror_even(kc, kz - 1)) + wval(g, wa, NULL, mirror_even(kc, kz))) +
           WC2 * (-wval(g, wa, NULL, mirror_even(kc, kz - 2)) + wval(g, wa, NULL, mirror_even(kc, kz + 1))) +
           WC3 * (-wval(g, wa, NULL, mirror_even(kc, kz - 3)) + wval(g, wa, NULL, mirror_even(kc, kz + 2))) +
           WC4 * (-wval(g, wa, NULL, mirror_even(kc, kz - 4)) + wval(g, wa, NULL, mirror_even(kc, kz + 3)));
}

/* Rotated D- : src/prop3DAcoTTIDenQ_DEO2_FDTD.h:3171-3657 */
void wo_tti_dminus_rot(long fs, long nx, long ny, long nz, float idx, float idy, float idz,
                       const float *g1, const float *g2, const float *g3,
                       const float *sT, const float *cT, const float *sP, const float *cP,
                       float *o1, float *o2, float *o3, long nbx, long nby, long nbz, long nthread) {
    const long nynz = ny * nz;
    float *outs[3] = {o1, o2, o3};
    zero_annulus(nx, ny, nz, outs, 3);
    FOR_INTERIOR(nx, ny, nz, nbx, nby, nbz, nthread, {
        const float A1 = sten_minus_w(g1, cT, cP, k, nynz);
        const float B1 = sten_minus_w(g1, cT, sP, k, nz);
        const float C1 = sten_minus_w(g1, sT, NULL, k, 1);
        const float A2 = sten_minus_w(g2, sP, NULL, k, nynz);
        const float B2 = sten_minus_w(g2, cP, NULL, k, nz);
        const float A3 = sten_minus_w(g3, sT, cP, k, nynz);
        const float B3 = sten_minus_w(g3, sT, sP, k, nz);
        const float C3 = sten_minus_w(g3, cT, NULL, k, 1);
        o1[k] = idx * A1 + idy * B1 - idz * C1;
        o2[k] = -idx * A2 + idy * B2;
        o3[k] = idx * A3 + idy * B3 + idz * C3;
    });
    if (fs) {
        FOR_FS_ROWS(nx, ny, nz, nthread, {
            if (kz == 0) {
                o1[k] = 0;
                o2[k] = 0;
                o3[k] = 0;
            } else {
                const float A1 = sten_minus_w(g1, cT, cP, k, nynz);
                const float B1 = sten_minus_w(g1, cT, sP, k, nz);
                const float C1 = sten_minus_w_fs(g1, sT, kc, kz);
                const float A2 = sten_minus_w(g2, sP, NULL, k, nynz);
                const float B2 = sten_minus_w(g2, cP, NULL, k, nz);
                const float A3 = sten_minus_w(g3, sT, cP, k, nynz);
                const float B3 = sten_minus_w(g3, sT, sP, k, nz);
                const float C3 = sten_minus_w_fs(g3, cT, kc, kz);
                o1[k] = idx * A1 + idy * B1 - idz * C1;
                o2[k] = -idx * A2 + idy * B2;
                o3[k] = idx * A3 + idy * B3 + idz * C3;
            }
        });
    }
}

/* ------------------------------------------------------------------------- */
/* Injection / extraction: src/spacetime.cc:132-204                            */
/* ------------------------------------------------------------------------- */

typedef struct {
    long iu;
    long ir;
    float c;
} wo_source_point32;

/* Points are visited in the given order; the reference's partitions keep equal keys
 * together and in order, so per-key accumulation order is the same. */
void wo_inject(float *field, const float *data, size_t it, size_t nt,
               const wo_source_point32 *pts, long npts) {
    for (long i = 0; i < npts; i++) {
        const wo_source_point32 p = pts[i];
        field[p.iu - 1] += p.c * data[(p.ir - 1) * nt + it - 1];
    }
}

void wo_extract(float *data, const float *field, const wo_source_point32 *pts, long npts,
                size_t it, size_t nt) {
    for (long i = 0; i < npts; i++) {
        const wo_source_point32 p = pts[i];
        data[(p.ir - 1) * nt + it - 1] += p.c * field[p.iu - 1];
    }
}

/* ------------------------------------------------------------------------- */
/* 8-tap time interpolation: src/spacetime.cc:4-127                           */
/* h is nh filters of 8 taps, row-major h[k*8+l].                             */
/* ------------------------------------------------------------------------- */

/* The reference visits coarse rows (0,[3,7]) (1,[2,7]) (2,[1,7]) (3..nd-5,[0,7])
 * (nd-4,[0,6]) (nd-3,[0,5]) (nd-2,[0,4]) in size_t arithmetic; rows whose index wraps are
 * skipped by the jk >= nm test and rows repeat when nd < 7.  Kept literally. */
static void interp_row(int fwd, const float *h, float *d, float *m, size_t nh, size_t nd, size_t nm,
                       size_t i, size_t lo, size_t hi) {
    const size_t j = i * nh, im3 = i - 3;
    for (size_t k = 0; k < nh; k++) {
        const size_t jk = j + k;
        if (jk >= nm) continue;
        for (size_t l = lo; l <= hi; l++) {
            if (im3 + l >= nd) continue; /* out of the trace: undefined in the reference (nd < 7) */
            if (fwd) d[im3 + l] += h[k * 8 + l] * m[jk];
            else m[jk] += h[k * 8 + l] * d[im3 + l];
        }
    }
}

static void interp_rows(int fwd, const float *h, float *d, float *m, size_t nh, size_t nd, size_t nm) {
    interp_row(fwd, h, d, m, nh, nd, nm, 0, 3, 7);
    interp_row(fwd, h, d, m, nh, nd, nm, 1, 2, 7);
    interp_row(fwd, h, d, m, nh, nd, nm, 2, 1, 7);
    for (size_t i = 3; i < nd - 4 && i * nh < nm; i++) interp_row(fwd, h, d, m, nh, nd, nm, i, 0, 7);
    interp_row(fwd, h, d, m, nh, nd, nm, nd - 4, 0, 6);
    interp_row(fwd, h, d, m, nh, nd, nm, nd - 3, 0, 5);
    interp_row(fwd, h, d, m, nh, nd, nm, nd - 2, 0, 4);
}

/* fine (m, nm samples) -> coarse (d, nd samples), per trace of nslow traces */
void wo_interp_forward(const float *h, float *d, const float *m, long nh, long nd, long nm, long nslow) {
    for (long s = 0; s < nslow; s++) {
        float *ds = d + s * nd;
        memset(ds, 0, sizeof(float) * nd);
        interp_rows(1, h, ds, (float *)(m + s * nm), (size_t)nh, (size_t)nd, (size_t)nm);
    }
}

/* coarse (d) -> fine (m) */
void wo_interp_adjoint(const float *h, float *m, const float *d, long nh, long nm, long nd, long nslow) {
    for (long s = 0; s < nslow; s++) {
        float *ms = m + s * nm;
        memset(ms, 0, sizeof(float) * nm);
        interp_rows(0, h, (float *)(d + s * nd), ms, (size_t)nh, (size_t)nd, (size_t)nm);
    }
}

/* ------------------------------------------------------------------------- */
/* Whole time steps (two-pass, as the reference timeStep() methods)            */
/* ------------------------------------------------------------------------- */

/* Iso timeStep (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:247-261) without the pointer swap:
 * writes p(t+dt) into pOld and L p into pSpace; caller swaps. tmp: 3*n floats scratch. */
void wo_iso_step(long fs, long nx, long ny, long nz, float idx, float idy, float idz, float dt,
                 const float *v, const float *b, const float *q, const float *pCur,
                 float *pSpace, float *pOld, float *tmp, long nbx, long nby, long nbz, long nthread) {
    const long n = nx * ny * nz;
    wo_iso_sandwich(fs, nx, ny, nz, idx, idy, idz, pCur, b, tmp, tmp + n, tmp + 2 * n, nbx, nby, nbz, nthread);
    wo_iso_update(fs, nx, ny, nz, idx, idy, idz, dt, tmp, tmp + n, tmp + 2 * n, v, b, q, pCur,
                  pSpace, pOld, nbx, nby, nbz, nthread);
}

/* VTI timeStep / timeStepLinear (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:326-376). tmp: 6*n. */
void wo_vti_step(long fs, long nx, long ny, long nz, float idx, float idy, float idz, float dt,
                 const float *v, const float *eps, const float *eta, const float *b, const float *f,
                 const float *q, const float *pCur, const float *mCur,
                 float *pSpace, float *mSpace, float *pOld, float *mOld, float *tmp,
                 long nbx, long nby, long nbz, long nthread) {
    const long n = nx * ny * nz;
    wo_vti_sandwich(fs, nx, ny, nz, idx, idy, idz, pCur, mCur, eps, eta, f, b,
                    tmp, tmp + n, tmp + 2 * n, tmp + 3 * n, tmp + 4 * n, tmp + 5 * n, nbx, nby, nbz, nthread);
    wo_vti_update(fs, nx, ny, nz, idx, idy, idz, dt, tmp, tmp + n, tmp + 2 * n, tmp + 3 * n,
                  tmp + 4 * n, tmp + 5 * n, v, b, q, pCur, mCur, pSpace, mSpace, pOld, mOld,
                  nbx, nby, nbz, nthread);
}

/* TTI timeStep (src/prop3DAcoTTIDenQ_DEO2_FDTD.h:359-379). tmp: 6*n. */
void wo_tti_step(long fs, long nx, long ny, long nz, float idx, float idy, float idz, float dt,
                 const float *v, const float *eps, const float *eta, const float *b, const float *f,
                 const float *q, const float *sT, const float *cT, const float *sP, const float *cP,
                 const float *pCur, const float *mCur,
                 float *pSpace, float *mSpace, float *pOld, float *mOld, float *tmp,
                 long nbx, long nby, long nbz, long nthread) {
    const long n = nx * ny * nz;
    wo_tti_sandwich(fs, nx, ny, nz, idx, idy, idz, pCur, mCur, sT, cT, sP, cP, eps, eta, f, b,
                    tmp, tmp + n, tmp + 2 * n, tmp + 3 * n, tmp + 4 * n, tmp + 5 * n, nbx, nby, nbz, nthread);
    wo_vti_update(fs, nx, ny, nz, idx, idy, idz, dt, tmp, tmp + n, tmp + 2 * n, tmp + 3 * n,
                  tmp + 4 * n, tmp + 5 * n, v, b, q, pCur, mCur, pSpace, mSpace, pOld, mOld,
                  nbx, nby, nbz, nthread);
}

int wo_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------- */
/* 2D: src/absorb.h:23-75, src/propagatorStaticFunctions.h:10-311,              */
/*     src/prop2DAcoIsoDenQ_DEO2_FDTD.h:674-1045                                */
/* Arrays (nx, nz), k = kx*nz + kz.  Outputs are zeroed whole, then the interior */
/* [4,nx-4) x [4,nz-4) and, under a free surface, the rows kz = 0..3.           */
/* ------------------------------------------------------------------------- */

/* Returns 0, or 1 when freqQ < FLT_EPSILON (the reference exits there, src/absorb.h:34-39). */
int wo2_setup_q(long fs, long nx, long nz, long nsponge, float dt, float freqQ, float qMin, float qInterior,
                float *q) {
    if (freqQ < FLT_EPSILON) return 1;
    if ((qMin < FLT_EPSILON) && (qInterior < FLT_EPSILON)) {
        long nzero = imin(2, nx * nz); /* sizeof(pointer) bytes (src/absorb.h:44) */
        for (long i = 0; i < nzero; i++) q[i] = 0;
        return 0;
    }
    float *qprof = (float *)malloc(sizeof(float) * (nsponge > 0 ? nsponge : 1));
    /* the reference is C++: log(qMin) on a float picks the float overload (logf), and the
     * result is widened to double for the ramp (src/absorb.h:47-48) */
    const double lqmin = (double)logf(qMin);
    const double lqmax = (double)logf(qInterior);
    for (long ks = 0; ks < nsponge; ks++) {
        const double dk = (double)(ks) / (double)(nsponge - 1);
        const double lq = lqmin + dk * (lqmax - lqmin);
        qprof[ks] = expf((float)lq); /* 2D: the ramp in double, exp in float (src/absorb.h:49-55) */
    }
    for (long kx = 0; kx < nx; kx++)
        for (long kz = 0; kz < nz; kz++) {
            const long ksx = imin(kx, nx - 1 - kx);
            const long ksz = fs ? (nz - 1 - kz) : imin(kz, nz - 1 - kz);
            const long ks = imin(ksx, ksz);
            /* dt * 2.0 is double here (src/absorb.h:67-71) */
            q[kx * nz + kz] = (float)((double)dt * 2.0 * WPI * (double)freqQ / (double)qInterior);
            if (ks < nsponge) q[kx * nz + kz] = (float)((double)dt * 2.0 * WPI * (double)freqQ / (double)qprof[ks]);
        }
    free(qprof);
    return 0;
}

#define FOR2_INTERIOR(nx, nz, ...)                                        \
    do {                                                                  \
        _Pragma("omp parallel for schedule(static)")                      \
        for (long kx = 4; kx < (nx) - 4; kx++)                            \
            for (long kz = 4; kz < (nz) - 4; kz++) {                      \
                const long k = kx * (nz) + kz;                            \
                __VA_ARGS__                                               \
            }                                                             \
    } while (0)

#define FOR2_FS_ROWS(nx, nz, ...)                                         \
    do {                                                                  \
        _Pragma("omp parallel for schedule(static)")                      \
        for (long kx = 4; kx < (nx) - 4; kx++)                            \
            for (long kz = 0; kz < 4; kz++) {                             \
                const long k = kx * (nz) + kz;                            \
                const long kc = kx * (nz);                                \
                __VA_ARGS__                                               \
            }                                                             \
    } while (0)

void wo2_dplus(long fs, long nx, long nz, float idx, float idz, const float *inX, const float *inZ, float *outX,
               float *outZ) {
    memset(outX, 0, sizeof(float) * nx * nz);
    memset(outZ, 0, sizeof(float) * nx * nz);
    FOR2_INTERIOR(nx, nz, {
        outX[k] = idx * sten_plus(inX, k, nz);
        outZ[k] = idz * sten_plus(inZ, k, 1);
    });
    if (fs) {
        FOR2_FS_ROWS(nx, nz, {
            outX[k] = kz == 0 ? 0.0f : idx * sten_plus(inX, k, nz);
            outZ[k] = idz * sten_plus_fs(inZ + kc, kz);
        });
    }
}

void wo2_dminus(long fs, long nx, long nz, float idx, float idz, const float *inX, const float *inZ, float *outX,
                float *outZ) {
    memset(outX, 0, sizeof(float) * nx * nz);
    memset(outZ, 0, sizeof(float) * nx * nz);
    FOR2_INTERIOR(nx, nz, {
        outX[k] = idx * sten_minus(inX, k, nz);
        outZ[k] = idz * sten_minus(inZ, k, 1);
    });
    if (fs) {
        FOR2_FS_ROWS(nx, nz, {
            if (kz == 0) {
                outX[k] = 0;
                outZ[k] = 0;
            } else {
                outX[k] = idx * sten_minus(inX, k, nz);
                outZ[k] = idz * sten_minus_fs(inZ + kc, kz);
            }
        });
    }
}

void wo2_iso_sandwich(long fs, long nx, long nz, float idx, float idz, const float *p, const float *b, float *tX,
                      float *tZ) {
    memset(tX, 0, sizeof(float) * nx * nz);
    memset(tZ, 0, sizeof(float) * nx * nz);
    FOR2_INTERIOR(nx, nz, {
        const float dPx = idx * sten_plus(p, k, nz);
        const float dPz = idz * sten_plus(p, k, 1);
        const float B = b[k];
        tX[k] = B * dPx;
        tZ[k] = B * dPz;
    });
    if (fs) {
        FOR2_FS_ROWS(nx, nz, {
            const float dPx = kz == 0 ? 0.0f : idx * sten_plus(p, k, nz);
            const float dPz = idz * sten_plus_fs(p + kc, kz);
            const float B = b[k];
            tX[k] = B * dPx;
            tZ[k] = B * dPz;
        });
    }
}

void wo2_iso_update(long fs, long nx, long nz, float idx, float idz, float dt, const float *tX, const float *tZ,
                    const float *v, const float *b, const float *q, const float *pCur, float *pSpace, float *pOld) {
    const float dt2 = dt * dt;
    memset(pSpace, 0, sizeof(float) * nx * nz);
    FOR2_INTERIOR(nx, nz, {
        const float dPx = idx * sten_minus(tX, k, nz);
        const float dPz = idz * sten_minus(tZ, k, 1);
        const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
        pOld[k] = dt2V2_B * (dPx + dPz) - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
        pSpace[k] = dPx + dPz;
    });
    if (fs) {
        FOR2_FS_ROWS(nx, nz, {
            float dPx = 0, dPz = 0;
            if (kz > 0) {
                dPx = idx * sten_minus(tX, k, nz);
                dPz = idz * sten_minus_fs(tZ + kc, kz);
            }
            const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
            pOld[k] = dt2V2_B * (dPx + dPz) - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
            pSpace[k] = dPx + dPz;
        });
    }
}

/* ------------------------------------------------------------------------- */
/* 2D VTI: src/prop2DAcoVTIDenQ_DEO2_FDTD.h:686-1477                           */
/* 2D TTI: src/prop2DAcoTTIDenQ_DEO2_FDTD.h:667-1717                           */
/* ------------------------------------------------------------------------- */

/* The 2D material sandwich (VTI :785-793; TTI :775-783 on the rotated derivatives).  Unlike
 * the 3D text, the coupling coefficient multiplies left to right: B * F * A * sqrt(1 - A*A). */
static inline void aniso2_material(float dPx, float dPz, float dMx, float dMz, float eps, float eta, float f,
                                   float b, float *tPx, float *tPz, float *tMx, float *tMz) {
    const float E = 1 + 2 * eps;
    const float A = eta;
    const float F = f;
    const float B = b;
    *tPx = B * E * dPx;
    *tPz = B * (1 - F * A * A) * dPz + B * F * A * sqrtf(1 - A * A) * dMz;
    *tMx = B * (1 - F) * dMx;
    *tMz = B * F * A * sqrtf(1 - A * A) * dPz + B * (1 - F + F * A * A) * dMz;
}

void wo2_vti_sandwich(long fs, long nx, long nz, float idx, float idz, const float *p, const float *m,
                      const float *eps, const float *eta, const float *f, const float *b, float *tPx, float *tPz,
                      float *tMx, float *tMz) {
    float *outs[4] = {tPx, tPz, tMx, tMz};
    for (int a = 0; a < 4; a++) memset(outs[a], 0, sizeof(float) * nx * nz);
    FOR2_INTERIOR(nx, nz, {
        const float dPx = idx * sten_plus(p, k, nz);
        const float dPz = idz * sten_plus(p, k, 1);
        const float dMx = idx * sten_plus(m, k, nz);
        const float dMz = idz * sten_plus(m, k, 1);
        aniso2_material(dPx, dPz, dMx, dMz, eps[k], eta[k], f[k], b[k], &tPx[k], &tPz[k], &tMx[k], &tMz[k]);
    });
    if (fs) {
        FOR2_FS_ROWS(nx, nz, {
            const float dPx = kz == 0 ? 0.0f : idx * sten_plus(p, k, nz);
            const float dPz = idz * sten_plus_fs(p + kc, kz);
            const float dMx = kz == 0 ? 0.0f : idx * sten_plus(m, k, nz);
            const float dMz = idz * sten_plus_fs(m + kc, kz);
            aniso2_material(dPx, dPz, dMx, dMz, eps[k], eta[k], f[k], b[k], &tPx[k], &tPz[k], &tMx[k], &tMz[k]);
        });
    }
}

/* P/M D- and leapfrog update (VTI :1032-1259 nonlinear, :1261-1474 linear).  pSpace/mSpace
 * NULL selects the linear variant, which neither clears nor writes them. */
void wo2_vti_update(long fs, long nx, long nz, float idx, float idz, float dt, const float *tPx, const float *tPz,
                    const float *tMx, const float *tMz, const float *v, const float *b, const float *q,
                    const float *pCur, const float *mCur, float *pSpace, float *mSpace, float *pOld, float *mOld) {
    const float dt2 = dt * dt;
    if (pSpace) {
        memset(pSpace, 0, sizeof(float) * nx * nz);
        memset(mSpace, 0, sizeof(float) * nx * nz);
    }
    FOR2_INTERIOR(nx, nz, {
        const float dPx = idx * sten_minus(tPx, k, nz);
        const float dPz = idz * sten_minus(tPz, k, 1);
        const float dMx = idx * sten_minus(tMx, k, nz);
        const float dMz = idz * sten_minus(tMz, k, 1);
        const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
        const float LP = dPx + dPz;
        const float LM = dMx + dMz;
        if (pSpace) { pSpace[k] = LP; mSpace[k] = LM; }
        pOld[k] = dt2V2_B * LP - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
        mOld[k] = dt2V2_B * LM - q[k] * (mCur[k] - mOld[k]) - mOld[k] + 2 * mCur[k];
    });
    if (fs) {
        FOR2_FS_ROWS(nx, nz, {
            float dPx = 0, dPz = 0, dMx = 0, dMz = 0;
            if (kz > 0) {
                dPx = idx * sten_minus(tPx, k, nz);
                dPz = idz * sten_minus_fs(tPz + kc, kz);
                dMx = idx * sten_minus(tMx, k, nz);
                dMz = idz * sten_minus_fs(tMz + kc, kz);
            }
            const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
            const float LP = dPx + dPz;
            const float LM = dMx + dMz;
            if (pSpace) { pSpace[k] = LP; mSpace[k] = LM; }
            pOld[k] = dt2V2_B * LP - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
            mOld[k] = dt2V2_B * LM - q[k] * (mCur[k] - mOld[k]) - mOld[k] + 2 * mCur[k];
        });
    }
}

/* TTI D+ with the rotation into the symmetry axis' frame (:1366-1528):
 * g1 = cos dx - sin dz, g3 = sin dx + cos dz; at the free surface g1 = -sin dz, g3 = cos dz. */
void wo2_tti_dplus(long fs, long nx, long nz, float idx, float idz, const float *g1, const float *g3,
                   const float *sT, const float *cT, float *o1, float *o3) {
    memset(o1, 0, sizeof(float) * nx * nz);
    memset(o3, 0, sizeof(float) * nx * nz);
    FOR2_INTERIOR(nx, nz, {
        const float dx = idx * sten_plus(g1, k, nz);
        const float dz = idz * sten_plus(g3, k, 1);
        o1[k] = cT[k] * dx - sT[k] * dz;
        o3[k] = sT[k] * dx + cT[k] * dz;
    });
    if (fs) {
        FOR2_FS_ROWS(nx, nz, {
            const float dz = idz * sten_plus_fs(g3 + kc, kz);
            if (kz == 0) {
                o1[k] = -sT[k] * dz;
                o3[k] = +cT[k] * dz;
            } else {
                const float dx = idx * sten_plus(g1, k, nz);
                o1[k] = cT[k] * dx - sT[k] * dz;
                o3[k] = sT[k] * dx + cT[k] * dz;
            }
        });
    }
}

/* TTI D- with the transposed rotation applied inside the differences (:1531-1716):
 * o1 = D-x(cos g1) - D-z(sin g1), o3 = D-x(sin g3) + D-z(cos g3); 0 at the free surface. */
static inline void tti2_minus(const float *g1, const float *g3, const float *sT, const float *cT, long k, long kc,
                              long kz, long nz, float idx, float idz, int fsrow, float *o1, float *o3) {
    const float A1 = sten_minus_w(g1, cT, NULL, k, nz);
    const float B1 = fsrow ? sten_minus_w_fs(g1, sT, kc, kz) : sten_minus_w(g1, sT, NULL, k, 1);
    const float A3 = sten_minus_w(g3, sT, NULL, k, nz);
    const float B3 = fsrow ? sten_minus_w_fs(g3, cT, kc, kz) : sten_minus_w(g3, cT, NULL, k, 1);
    *o1 = idx * A1 - idz * B1;
    *o3 = idx * A3 + idz * B3;
}

void wo2_tti_dminus(long fs, long nx, long nz, float idx, float idz, const float *g1, const float *g3,
                    const float *sT, const float *cT, float *o1, float *o3) {
    memset(o1, 0, sizeof(float) * nx * nz);
    memset(o3, 0, sizeof(float) * nx * nz);
    FOR2_INTERIOR(nx, nz, { tti2_minus(g1, g3, sT, cT, k, kx * nz, kz, nz, idx, idz, 0, &o1[k], &o3[k]); });
    if (fs) {
        FOR2_FS_ROWS(nx, nz, {
            if (kz == 0) {
                o1[k] = 0;
                o3[k] = 0;
            } else {
                tti2_minus(g1, g3, sT, cT, k, kc, kz, nz, idx, idz, 1, &o1[k], &o3[k]);
            }
        });
    }
}

/* TTI sandwich (:667-994): rotated D+ of P and M, then the material */
void wo2_tti_sandwich(long fs, long nx, long nz, float idx, float idz, const float *p, const float *m,
                      const float *sT, const float *cT, const float *eps, const float *eta, const float *f,
                      const float *b, float *tP1, float *tP3, float *tM1, float *tM3) {
    wo2_tti_dplus(fs, nx, nz, idx, idz, p, p, sT, cT, tP1, tP3);
    wo2_tti_dplus(fs, nx, nz, idx, idz, m, m, sT, cT, tM1, tM3);
    const long zlo = fs ? 0 : 4;
#pragma omp parallel for schedule(static)
    for (long kx = 4; kx < nx - 4; kx++)
        for (long kz = zlo; kz < nz - 4; kz++) {
            const long k = kx * nz + kz;
            aniso2_material(tP1[k], tP3[k], tM1[k], tM3[k], eps[k], eta[k], f[k], b[k], &tP1[k], &tP3[k], &tM1[k],
                            &tM3[k]);
        }
}

/* TTI D- and leapfrog update (:997-1363), nonlinear only (the 2D TTI has no linear step) */
void wo2_tti_update(long fs, long nx, long nz, float idx, float idz, float dt, const float *tP1, const float *tP3,
                    const float *tM1, const float *tM3, const float *sT, const float *cT, const float *v,
                    const float *b, const float *q, const float *pCur, const float *mCur, float *pSpace,
                    float *mSpace, float *pOld, float *mOld) {
    const float dt2 = dt * dt;
    memset(pSpace, 0, sizeof(float) * nx * nz);
    memset(mSpace, 0, sizeof(float) * nx * nz);
    const long zlo = fs ? 0 : 4;
#pragma omp parallel for schedule(static)
    for (long kx = 4; kx < nx - 4; kx++)
        for (long kz = zlo; kz < nz - 4; kz++) {
            const long k = kx * nz + kz;
            float dP1 = 0, dP3 = 0, dM1 = 0, dM3 = 0;
            if (!(fs && kz == 0)) {
                const int fsrow = fs && kz < 4;
                tti2_minus(tP1, tP3, sT, cT, k, kx * nz, kz, nz, idx, idz, fsrow, &dP1, &dP3);
                tti2_minus(tM1, tM3, sT, cT, k, kx * nz, kz, nz, idx, idz, fsrow, &dM1, &dM3);
            }
            const float dt2V2_B = dt2 * v[k] * v[k] / b[k];
            pSpace[k] = dP1 + dP3;
            mSpace[k] = dM1 + dM3;
            pOld[k] = dt2V2_B * pSpace[k] - q[k] * (pCur[k] - pOld[k]) - pOld[k] + 2 * pCur[k];
            mOld[k] = dt2V2_B * mSpace[k] - q[k] * (mCur[k] - mOld[k]) - mOld[k] + 2 * mCur[k];
        }
}

/* Flush-to-zero and denormals-are-zero on every thread of an nthread OpenMP team (MXCSR bits 15
 * and 6): the reference's timing loop runs with set_zero_subnormals(true)
 * (src/prop3DAcoTTIDenQ_DEO2_FDTD.jl:172).  For timing only; the parity tests keep IEEE
 * subnormals (results differ only where a value would be subnormal). */
void wo_set_ftz(long nthread, int on) {
#pragma omp parallel num_threads(nthread)
    {
        unsigned m = _mm_getcsr();
        m = on ? (m | 0x8040u) : (m & ~0x8040u);
        _mm_setcsr(m);
    }
    unsigned m = _mm_getcsr();
    _mm_setcsr(on ? (m | 0x8040u) : (m & ~0x8040u));
}
