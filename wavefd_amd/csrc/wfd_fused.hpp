// wfd_fused.hpp -- single-pass fused time step (D+ sandwich, D-, leapfrog with Q) for the
// three physics.  One kernel replaces the reference's two OpenMP passes
// (e.g. src/prop3DAcoTTIDenQ_DEO2_FDTD.h:359-379): the D+ temporaries never reach HBM.
//
// Decomposition ("2.5D", marching along x, the slowest axis):
//   * a workgroup owns an output tile of TZO z-cells x TYO y-cells and a chunk of CX x-planes;
//   * its 64 x (TYO+7) threads cover the tile plus the 4-cell halo (-4..+3) that the
//     backward difference D- needs, lanes along z (coalesced), one row of y per wavefront;
//   * every thread keeps, in registers, the last 8 x-planes of p (and m) of its column and
//     the last 8 x-planes of the x-temporaries; the temporaries along y and z are exchanged
//     through LDS;
//   * at iteration x a thread loads p(x+7), computes the full gradient and the material
//     sandwich at plane x+3 (D+y, D+z from an LDS copy of plane x+3 with a 7-cell border),
//     and the output threads then apply D- and the leapfrog update at plane x.
// Halo threads recompute the sandwich of their neighbours' cells, which is what lets TTI's
// rotated sandwich (full gradient at every temporary) fuse at all.
//
// Semantics match the operator kernels exactly: temporaries are 0 outside the update
// region (the reference's zeroed annulus), the free surface uses the odd mirror for D+ and
// the even mirror for D-, x/y derivatives vanish at kz=0 and L=0 there.  Expression order
// is unchanged, so results are bit-identical to the two-pass path.
#pragma once
#include "wfd_kernels.hpp"
#include "wfd_runtime.hpp"

namespace wfd {

constexpr int FZ = 64;        // threads along z per row (one wavefront)
constexpr int FZO = 56;       // output z cells per tile (lanes 4..59)
constexpr int FYO = 8;        // output y rows per tile
constexpr int FYE = FYO + 7;  // rows of threads (tile + halo)
constexpr int FLZ = FZ + 7;   // LDS plane width along z: lanes -3 .. 67 around the E lanes
constexpr int FLY = FYE + 7;  // LDS plane rows

template <int PHYS>
struct FusedArgs {
    Grid g;
    float dt2;
    long cx;  // x-planes per chunk
    const float *pCur, *mCur;
    float *pOld, *mOld, *pSpace, *mSpace;
    const float *v, *b, *q, *eps, *eta, *f, *sT, *cT, *sP, *cP;
};

__device__ __forceinline__ float sten8(const float *a) {
    // a[0..7] = values at offsets -3..+4 (D+) or -4..+3 (D-) of the centre pair a[3],a[4]
    return C1 * (-a[3] + a[4]) + C2 * (-a[2] + a[5]) + C3 * (-a[1] + a[6]) + C4 * (-a[0] + a[7]);
}

template <int PHYS>
__global__ void __launch_bounds__(FZ *FYE) k_fused_step(FusedArgs<PHYS> A) {
    constexpr int NF = PHYS == ISO ? 1 : 2;
    const Grid &g = A.g;
    const int l = threadIdx.x;  // lane
    const int r = threadIdx.y;  // row
    const long z0 = (long)blockIdx.x * FZO;
    const long y0 = (long)blockIdx.y * FYO;
    const long xs = (long)blockIdx.z * A.cx;
    const long xe = xs + A.cx < g.nx ? xs + A.cx : g.nx;
    const long z = z0 - 4 + l;
    const long y = y0 - 4 + r;
    const bool zin = z >= 0 && z < g.nz, yin = y >= 0 && y < g.ny;
    const bool col_ok = zin && yin;
    const long zlo = g.fs ? 0 : 4;
    const bool yzU = y >= 4 && y < g.ny - 4 && z >= zlo && z < g.nz - 4;  // (y,z) inside the update region
    const bool out_thread = l >= 4 && l < 4 + FZO && r >= 4 && r < 4 + FYO && col_ok;
    const long colk = y * g.nz + z;  // offset inside a plane (valid only when col_ok)

    __shared__ float sp[NF][FLY][FLZ];      // p (and m) at plane x+3, border 7
    __shared__ float soy[NF][FYE][FZ];      // y-temporaries at plane x, E rows
    __shared__ float soz[NF][FYO][FZ + 1];  // z-temporaries at plane x, output rows

    const float *fld[2] = {A.pCur, A.mCur};
    float pq[NF][8];   // p(x .. x+7) of this column
    float oxq[NF][8];  // x-temporaries at planes x-4 .. x+3
    float ry[NF][4], rz[NF][4];  // y/z temporaries of planes x .. x+3 (ring, slot = plane & 3)
    float bq[4];                 // b at planes x .. x+3 (slot = plane & 3)

    auto ld = [&](const float *a, long xx) -> float {
        return (col_ok && xx >= 0 && xx < g.nx) ? a[xx * g.nynz + colk] : 0.0f;
    };

#pragma unroll
    for (int fi = 0; fi < NF; fi++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            pq[fi][j] = 0.0f;
            oxq[fi][j] = 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            ry[fi][j] = 0.0f;
            rz[fi][j] = 0.0f;
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) bq[j] = 0.0f;
    // preload p(x0-7 .. x0-1) into slots 1..7 so the first iteration shifts them to 0..6
    const long xfirst = xs - 7;
#pragma unroll
    for (int fi = 0; fi < NF; fi++)
#pragma unroll
        for (int j = 1; j < 8; j++) pq[fi][j] = ld(fld[fi], xfirst + j - 1);

    for (long x = xfirst; x < xe; x++) {
        const long xp = x + 3;  // plane whose temporaries are produced this iteration
        // ---- shift queues, load p(x+7) ------------------------------------------------
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
#pragma unroll
            for (int j = 0; j < 7; j++) {
                pq[fi][j] = pq[fi][j + 1];
                oxq[fi][j] = oxq[fi][j + 1];
            }
            pq[fi][7] = ld(fld[fi], x + 7);
        }
        const bool xpU = xp >= 4 && xp < g.nx - 4;
        const bool inU = xpU && yzU;  // (x+3, y, z) in the update region
        // ---- stage plane x+3 of p (and m) in LDS with a 7-cell border -------------------
        __syncthreads();  // previous iteration's readers are done
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
            float val = pq[fi][3];
            if (g.fs && z < 0 && yin && xp >= 0 && xp < g.nx && -z < g.nz)
                val = -fld[fi][xp * g.nynz + y * g.nz - z];  // odd mirror above the surface
            sp[fi][r + 3][l + 3] = val;
        }
        for (int s = r * FZ + l; s < FLY * FLZ; s += FZ * FYE) {
            const int sy = s / FLZ, sz = s - sy * FLZ;
            if (sy >= 3 && sy < 3 + FYE && sz >= 3 && sz < 3 + FZ) continue;  // owned by a thread
            const long yy = y0 - 7 + sy, zz = z0 - 7 + sz;
            float vals[NF];
#pragma unroll
            for (int fi = 0; fi < NF; fi++) vals[fi] = 0.0f;
            if (yy >= 0 && yy < g.ny && xp >= 0 && xp < g.nx) {
                if (zz >= 0 && zz < g.nz) {
#pragma unroll
                    for (int fi = 0; fi < NF; fi++) vals[fi] = fld[fi][xp * g.nynz + yy * g.nz + zz];
                } else if (g.fs && zz < 0 && -zz < g.nz) {
#pragma unroll
                    for (int fi = 0; fi < NF; fi++) vals[fi] = -fld[fi][xp * g.nynz + yy * g.nz - zz];
                }
            }
#pragma unroll
            for (int fi = 0; fi < NF; fi++) sp[fi][sy][sz] = vals[fi];
        }
        __syncthreads();
        // ---- gradient + material sandwich at (x+3, y, z) --------------------------------
        float ox[NF], oy[NF], oz[NF];
#pragma unroll
        for (int fi = 0; fi < NF; fi++) ox[fi] = oy[fi] = oz[fi] = 0.0f;
        if (inU) {
            float dx[NF], dy[NF], dz[NF];
#pragma unroll
            for (int fi = 0; fi < NF; fi++) {
                float a[8];
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = sp[fi][r + j][l + 3];  // rows y-3 .. y+4
                dy[fi] = g.idy * sten8(a);
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = sp[fi][r + 3][l + j];  // lanes z-3 .. z+4
                dz[fi] = g.idz * sten8(a);
                dx[fi] = g.idx * sten8(pq[fi]);
                if (z == 0) {  // free-surface row: x/y derivatives vanish
                    dx[fi] = 0.0f;
                    dy[fi] = 0.0f;
                }
            }
            const long k = xp * g.nynz + colk;
            if constexpr (PHYS == ISO) {
                const float B = A.b[k];
                bq[xp & 3] = B;
                ox[0] = B * dx[0];
                oy[0] = B * dy[0];
                oz[0] = B * dz[0];
            } else if constexpr (PHYS == VTI) {
                const float B = A.b[k];
                bq[xp & 3] = B;
                vti_material(dx[0], dy[0], dz[0], dx[1], dy[1], dz[1], A.eps[k], A.eta[k], A.f[k], B, ox[0], oy[0],
                             oz[0], ox[1], oy[1], oz[1]);
            } else {
                const float B = A.b[k];
                bq[xp & 3] = B;
                tti_material(dx[0], dy[0], dz[0], dx[1], dy[1], dz[1], A.sT[k], A.cT[k], A.sP[k], A.cP[k], A.eps[k],
                             A.eta[k], A.f[k], B, ox[0], oy[0], oz[0], ox[1], oy[1], oz[1]);
            }
        }
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
            oxq[fi][7] = ox[fi];
            ry[fi][xp & 3] = oy[fi];
            rz[fi][xp & 3] = oz[fi];
        }
        if (x < xs) continue;  // warm-up iterations produce temporaries only
        // ---- exchange plane-x temporaries along y and z ------------------------------
        // (soy/soz of the previous iteration were last read before this iteration's
        // first barrier, so they can be overwritten now)
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
            soy[fi][r][l] = ry[fi][x & 3];
            if (r >= 4 && r < 4 + FYO) soz[fi][r - 4][l] = rz[fi][x & 3];
        }
        __syncthreads();
        if (!out_thread) continue;
        const long k = x * g.nynz + colk;
        const bool xU = x >= 4 && x < g.nx - 4;
        if (!(xU && yzU)) {
            // annulus: the spatial output is zeroed, the wavefield is not updated
            if (A.pSpace) {
                A.pSpace[k] = 0.0f;
                if (NF == 2) A.mSpace[k] = 0.0f;
            }
            continue;
        }
        float L[NF];
#pragma unroll
        for (int fi = 0; fi < NF; fi++) {
            if (z == 0) {
                L[fi] = 0.0f + 0.0f + 0.0f;
                continue;
            }
            float a[8];
            const float dxm = g.idx * sten8(oxq[fi]);
#pragma unroll
            for (int j = 0; j < 8; j++) a[j] = soy[fi][r - 4 + j][l];  // rows y-4 .. y+3
            const float dym = g.idy * sten8(a);
            if (g.fs && z < 4) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const long zz = z - 4 + j;           // even mirror: t[-j] = t[j-1]
                    const long zm = zz >= 0 ? zz : -zz - 1;
                    a[j] = soz[fi][r - 4][(int)(zm - (z0 - 4))];
                }
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = soz[fi][r - 4][l - 4 + j];  // lanes z-4 .. z+3
            }
            const float dzm = g.idz * sten8(a);
            L[fi] = dxm + dym + dzm;
        }
        const float V = A.v[k];
        const float dt2V2_B = A.dt2 * V * V / bq[x & 3];
        const float qk = A.q[k];
        const float pc = pq[0][0];
        A.pOld[k] = leapfrog(dt2V2_B, L[0], qk, pc, A.pOld[k]);
        if (A.pSpace) A.pSpace[k] = L[0];
        if constexpr (NF == 2) {
            const float mc = pq[1][0];
            A.mOld[k] = leapfrog(dt2V2_B, L[1], qk, mc, A.mOld[k]);
            if (A.mSpace) A.mSpace[k] = L[1];
        }
    }
}

static inline long fused_cx(const Prop *p) {
    if (const char *e = std::getenv("WAVEFD_FUSED_CX")) {
        long v = std::atol(e);
        if (v > 0) return v;
    }
    // enough workgroups to fill 256 CUs several times over, without paying the 7-plane
    // warm-up of each chunk too often
    const long tiles = ((p->nz + FZO - 1) / FZO) * ((p->ny + FYO - 1) / FYO);
    long cx = p->nx;
    while (cx > 32 && tiles * ((p->nx + cx - 1) / cx) < 2048) cx = (cx + 1) / 2;
    return cx;
}

static inline bool step_fused(Prop *p, bool linear) {
    const Field *f = p->f;
    const long cx = fused_cx(p);
    dim3 grid((unsigned)((p->nz + FZO - 1) / FZO), (unsigned)((p->ny + FYO - 1) / FYO),
              (unsigned)((p->nx + cx - 1) / cx));
    dim3 block(FZ, FYE, 1);
#define WFD_FILL_ARGS(A)                                                            \
    A.g = p->g;                                                                     \
    A.dt2 = p->dt * p->dt;                                                          \
    A.cx = cx;                                                                      \
    A.pCur = f[F_PCUR].d;                                                           \
    A.pOld = f[F_POLD].d;                                                           \
    A.pSpace = linear ? nullptr : f[F_PSPACE].d;                                    \
    A.mCur = f[F_MCUR].d;                                                           \
    A.mOld = f[F_MOLD].d;                                                           \
    A.mSpace = linear ? nullptr : f[F_MSPACE].d;                                    \
    A.v = f[F_V].d;                                                                 \
    A.b = f[F_B].d;                                                                 \
    A.q = f[F_Q].d;                                                                 \
    A.eps = f[F_EPS].d;                                                             \
    A.eta = f[F_ETA].d;                                                             \
    A.f = f[F_F].d;                                                                 \
    A.sT = f[F_ST].d;                                                               \
    A.cT = f[F_CT].d;                                                               \
    A.sP = f[F_SP].d;                                                               \
    A.cP = f[F_CP].d;
    if (p->phys == ISO) {
        FusedArgs<ISO> A;
        WFD_FILL_ARGS(A);
        k_fused_step<ISO><<<grid, block, 0, p->stream>>>(A);
    } else if (p->phys == VTI) {
        FusedArgs<VTI> A;
        WFD_FILL_ARGS(A);
        k_fused_step<VTI><<<grid, block, 0, p->stream>>>(A);
    } else {
        FusedArgs<TTI> A;
        WFD_FILL_ARGS(A);
        k_fused_step<TTI><<<grid, block, 0, p->stream>>>(A);
    }
#undef WFD_FILL_ARGS
    WFD_LAUNCHED();
    return true;
}

}  // namespace wfd
