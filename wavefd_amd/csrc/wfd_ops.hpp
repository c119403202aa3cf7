// wfd_ops.hpp -- the operations behind each C-ABI entry point, written once for the three
// physics.  Every function brackets its device work with Prop::begin/end so the
// reference's host-visible semantics hold in COHERENT mode (see wfd_runtime.hpp).
#pragma once
#include <hipfft/hipfft.h>

#include "wfd_2d.hpp"
#include "wfd_fused.hpp"
#include "wfd_runtime.hpp"

namespace wfd {

// ---------------------------------------------------------------------------------------
// time step
// ---------------------------------------------------------------------------------------

// Two-pass step exactly as the reference structures it: sandwich into temporaries, then
// D- + update (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:247-261, src/prop3DAcoVTIDenQ_DEO2_FDTD.h:326-376,
// src/prop3DAcoTTIDenQ_DEO2_FDTD.h:359-379).  Kept as the cross-check of the fused kernels.
static inline bool step_two_pass(Prop *p, bool linear) {
    const Grid &g = p->g;
    const float dt2 = p->dt * p->dt;
    float *t[6];
    const int nt = p->phys == ISO ? 3 : 6;
    for (int i = 0; i < nt; i++)
        if (!(t[i] = p->scratch(i))) return false;
    Field *f = p->f;
    if (p->phys == ISO) {
        k_iso_sandwich<<<op_grid(g), op_block(), 0, p->stream>>>(g, f[F_PCUR].d, f[F_B].d, t[0], t[1], t[2]);
        WFD_LAUNCHED();
        k_iso_update<<<op_grid(g), op_block(), 0, p->stream>>>(g, dt2, t[0], t[1], t[2], f[F_V].d, f[F_B].d,
                                                               f[F_Q].d, f[F_PCUR].d,
                                                               linear ? nullptr : f[F_PSPACE].d, f[F_POLD].d);
        WFD_LAUNCHED();
        return true;
    }
    if (p->phys == VTI) {
        k_vti_sandwich<<<op_grid(g), op_block(), 0, p->stream>>>(g, f[F_PCUR].d, f[F_MCUR].d, f[F_EPS].d, f[F_ETA].d,
                                                                 f[F_F].d, f[F_B].d, t[0], t[1], t[2], t[3], t[4],
                                                                 t[5]);
    } else {
        k_tti_sandwich<<<op_grid(g), op_block(), 0, p->stream>>>(
            g, f[F_PCUR].d, f[F_MCUR].d, f[F_ST].d, f[F_CT].d, f[F_SP].d, f[F_CP].d, f[F_EPS].d, f[F_ETA].d, f[F_F].d,
            f[F_B].d, t[0], t[1], t[2], t[3], t[4], t[5]);
    }
    WFD_LAUNCHED();
    k_vti_update<<<op_grid(g), op_block(), 0, p->stream>>>(
        g, dt2, t[0], t[1], t[2], t[3], t[4], t[5], f[F_V].d, f[F_B].d, f[F_Q].d, f[F_PCUR].d, f[F_MCUR].d,
        linear ? nullptr : f[F_PSPACE].d, linear ? nullptr : f[F_MSPACE].d, f[F_POLD].d, f[F_MOLD].d);
    WFD_LAUNCHED();
    return true;
}

static inline bool step_device(Prop *p, bool linear) {
    bool ok = p->dim == 2 ? step_2d(p, linear) : (p->fused ? step_fused(p, linear) : step_two_pass(p, linear));
    if (!ok) return false;
    p->swap_p();
    return true;
}

static inline Mask step_writes(Prop *p, bool linear) {
    // after the swap the new field is pCur; the previous pCur became pOld (unchanged data)
    Mask m = bit(F_PCUR);
    if (!linear) m |= bit(F_PSPACE);
    if (p->phys != ISO) {
        m |= bit(F_MCUR);
        if (!linear) m |= bit(F_MSPACE);
    }
    return m;
}

static inline bool time_step(Prop *p, bool linear) {
    if (!p->begin(p->all_mask())) return false;
    if (!step_device(p, linear)) return false;
    return p->end(step_writes(p, linear));
}

// step, then inject the registered sources at time index it into the new pCur (and mCur),
// then extract the registered receivers at it -- the loop of test/spacetime.jl:486-492
// plus extraction, as one call.
static inline bool time_step_inject_extract(Prop *p, long it, bool linear) {
    if (!registered_it_ok(p, it, it)) return false;  // before the step: nothing changes on error
    if (!p->begin(p->all_mask())) return false;
    if (!step_device(p, linear)) return false;
    if (!inject_registered(p, it)) return false;
    if (!extract_registered(p, it)) return false;
    return p->end(step_writes(p, linear));
}

// nsteps calls of time_step_inject_extract (time indices it0 .. it0+nsteps-1) as a HIP graph:
// two steps (the period of the pCur/pOld swap) are captured once on the handle's stream, with
// the injection and extraction reading the time index from a device counter that the graph
// advances, and the graph is launched nsteps/2 times; an odd last step runs as a call.  The
// kernels, their arguments and their order are those of the calls one by one.  The handle must
// be in DEVICE mode (the coherent mode copies host arrays inside every call) and use the fused
// step (the two-pass step allocates its scratch on first use).  Returns when the steps are done.
static inline bool time_loop_graph(Prop *p, long it0, long nsteps) {
    if (nsteps <= 0) return true;
    if (p->coherence != DEVICE) {
        set_error("TimeStepInjectExtractGraph: the handle must be in DEVICE mode");
        return false;
    }
    if (p->dim != 3 || !p->fused) {
        set_error("TimeStepInjectExtractGraph: needs the fused 3D step");
        return false;
    }
    if (!registered_it_ok(p, it0, it0 + nsteps - 1)) return false;
    const long npairs = nsteps / 2;
    if (npairs > 0) {
        if (!p->it_dev) WFD_TRY(hipMalloc(&p->it_dev, sizeof(long)));
        k_set_i64<<<1, 1, 0, p->stream>>>(p->it_dev, it0);
        WFD_LAUNCHED();
        WFD_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
        long swaps = 0;  // step_device swaps the pCur/pOld (mCur/mOld) buffers on the host as it records
        bool ok = true;
        for (int k = 0; k < 2 && ok; k++) {
            ok = step_device(p, false);
            if (!ok) break;
            swaps++;
            ok = inject_registered(p, 0, p->it_dev) && extract_registered(p, 0, p->it_dev);
            if (ok) {
                k_inc_i64<<<1, 1, 0, p->stream>>>(p->it_dev);
                ok = hipGetLastError() == hipSuccess;
            }
        }
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(p->stream, &graph);
        hipGraphExec_t exec = nullptr;
        ok = ok && ec == hipSuccess && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
        for (long r = 0; r < npairs && ok; r++) ok = hipGraphLaunch(exec, p->stream) == hipSuccess;
        if (ok) ok = hipStreamSynchronize(p->stream) == hipSuccess;
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (!ok) {
            // the buffers swapped an even number of times after a complete capture; if the
            // capture broke off, undo the swaps it recorded
            if (swaps & 1) p->swap_p();
            if (g_last_error.empty()) set_error("TimeStepInjectExtractGraph: graph capture or launch failed");
            (void)hipGetLastError();
            return false;
        }
    }
    if (nsteps & 1) {
        const long it = it0 + nsteps - 1;
        if (!step_device(p, false) || !inject_registered(p, it) || !extract_registered(p, it)) return false;
        WFD_TRY(hipStreamSynchronize(p->stream));
    }
    return true;
}

// the registered injection / extraction alone, between other calls (Born loops: step, Born
// source, extract; adjoint: step, inject the adjoint source, image)
static inline bool inject_registered_call(Prop *p, long it) {
    const Mask m = bit(F_PCUR) | (p->phys != ISO ? bit(F_MCUR) : 0);
    if (!p->begin(m)) return false;
    if (!inject_registered(p, it)) return false;
    return p->end(m);
}

static inline bool extract_registered_call(Prop *p, long it) {
    const Mask m = bit(F_PCUR) | (p->phys != ISO ? bit(F_MCUR) : 0);
    if (!p->begin(m)) return false;
    if (!extract_registered(p, it)) return false;
    return p->end(0);
}

// ---------------------------------------------------------------------------------------
// Born source injection
// ---------------------------------------------------------------------------------------

#define WFD_ARG(ptr, slot, in, out, a) \
    if (!p->arg((ptr), (slot), (in), (out), (a))) return false

static inline bool iso_inj_v(Prop *p, float *dV, float *DP) {
    if (!p->begin(bit(F_V) | bit(F_B) | bit(F_PCUR))) return false;
    Prop::Arg a[2];
    WFD_ARG(dV, 0, true, false, a[0]);
    WFD_ARG(DP, 1, true, false, a[1]);
    k_iso_inj_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, p->f[F_V].d, p->f[F_B].d, a[0].dev, a[1].dev,
                                                           p->f[F_PCUR].d);
    WFD_LAUNCHED();
    if (!p->finish_args(a, 2)) return false;
    return p->end(bit(F_PCUR));
}

// VB (dV != null) and B (dV == null) sources (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:351-498)
static inline bool iso_inj_vb(Prop *p, float *dV, float *dB, float *P, float *DP) {
    if (!p->begin(bit(F_V) | bit(F_B) | bit(F_PCUR))) return false;
    Prop::Arg a[4];
    int na = 0;
    if (dV) WFD_ARG(dV, 0, true, false, a[na++]);
    Prop::Arg &adB = a[na];
    WFD_ARG(dB, 1, true, false, a[na++]);
    Prop::Arg &aP = a[na];
    WFD_ARG(P, 2, true, false, a[na++]);
    Prop::Arg &aDP = a[na];
    WFD_ARG(DP, 3, true, false, a[na++]);
    float *t[6];
    for (int i = 0; i < 6; i++)
        if (!(t[i] = p->scratch(i))) return false;
    if (!dplus(p, aP.dev, t[0], t[1], t[2])) return false;
    k_scale3<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, adB.dev, F3{t[0], t[1], t[2]}, W3{t[3], t[4], t[5]});
    WFD_LAUNCHED();
    if (!dminus(p, t[3], t[4], t[5], t[0], t[1], t[2])) return false;
    k_iso_inj_vb<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, p->f[F_V].d, p->f[F_B].d,
                                                            dV ? a[0].dev : nullptr, adB.dev, aDP.dev,
                                                            F3{t[0], t[1], t[2]}, p->f[F_PCUR].d);
    WFD_LAUNCHED();
    if (!p->finish_args(a, na)) return false;
    return p->end(bit(F_PCUR));
}

static inline bool vti_inj_v(Prop *p, float *dV, float *DP, float *DM) {
    if (!p->begin(bit(F_V) | bit(F_PCUR) | bit(F_MCUR))) return false;
    Prop::Arg a[3];
    WFD_ARG(dV, 0, true, false, a[0]);
    WFD_ARG(DP, 1, true, false, a[1]);
    WFD_ARG(DM, 2, true, false, a[2]);
    k_vti_inj_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, p->f[F_V].d, a[0].dev, a[1].dev, a[2].dev,
                                                           p->f[F_PCUR].d, p->f[F_MCUR].d);
    WFD_LAUNCHED();
    if (!p->finish_args(a, 3)) return false;
    return p->end(bit(F_PCUR) | bit(F_MCUR));
}

// VEA source (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:457-551, src/prop3DAcoTTIDenQ_DEO2_FDTD.h:460-554)
static inline bool vea_inj(Prop *p, float *dV, float *dE, float *dA, float *P, float *M, float *DP, float *DM) {
    if (!p->begin(p->model_mask() | bit(F_PCUR) | bit(F_MCUR))) return false;
    Prop::Arg a[7];
    WFD_ARG(dV, 0, true, false, a[0]);
    WFD_ARG(dE, 1, true, false, a[1]);
    WFD_ARG(dA, 2, true, false, a[2]);
    WFD_ARG(P, 3, true, false, a[3]);
    WFD_ARG(M, 4, true, false, a[4]);
    WFD_ARG(DP, 5, true, false, a[5]);
    WFD_ARG(DM, 6, true, false, a[6]);
    float *t[12];
    for (int i = 0; i < 12; i++)
        if (!(t[i] = p->scratch(i))) return false;
    if (!rplus(p, a[3].dev, t[0], t[1], t[2])) return false;
    if (!rplus(p, a[4].dev, t[3], t[4], t[5])) return false;
    k_vea_inj_mid<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->f[F_B].d, p->f[F_ETA].d, p->f[F_F].d, a[1].dev,
                                                             a[2].dev, F3{t[0], t[1], t[2]}, F3{t[3], t[4], t[5]},
                                                             W3{t[6], t[7], t[8]}, W3{t[9], t[10], t[11]});
    WFD_LAUNCHED();
    if (!rminus(p, t[6], t[7], t[8], t[0], t[1], t[2])) return false;
    if (!rminus(p, t[9], t[10], t[11], t[3], t[4], t[5])) return false;
    k_vea_inj_final<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, p->f[F_V].d, p->f[F_B].d, a[0].dev,
                                                               a[5].dev, a[6].dev, F3{t[0], t[1], t[2]},
                                                               F3{t[3], t[4], t[5]}, p->f[F_PCUR].d, p->f[F_MCUR].d);
    WFD_LAUNCHED();
    if (!p->finish_args(a, 7)) return false;
    return p->end(bit(F_PCUR) | bit(F_MCUR));
}

// ---------------------------------------------------------------------------------------
// Born imaging conditions
// ---------------------------------------------------------------------------------------

static inline bool iso_acc_v(Prop *p, float *gV, float *DP) {
    if (!p->begin(bit(F_V) | bit(F_B) | bit(F_POLD))) return false;
    Prop::Arg a[2];
    WFD_ARG(gV, 0, true, true, a[0]);
    WFD_ARG(DP, 1, true, false, a[1]);
    k_iso_acc_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->f[F_V].d, p->f[F_B].d, a[1].dev, p->f[F_POLD].d,
                                                           a[0].dev);
    WFD_LAUNCHED();
    if (!p->finish_args(a, 2)) return false;
    return p->end(0);
}

// VB (gV != null) and B imaging (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:563-656)
static inline bool iso_acc_vb(Prop *p, float *gV, float *gB, float *P, float *DP) {
    if (!p->begin(bit(F_V) | bit(F_B) | bit(F_POLD))) return false;
    Prop::Arg a[4];
    int na = 0;
    if (gV) WFD_ARG(gV, 0, true, true, a[na++]);
    Prop::Arg &agB = a[na];
    WFD_ARG(gB, 1, true, true, a[na++]);
    Prop::Arg &aP = a[na];
    WFD_ARG(P, 2, true, false, a[na++]);
    Prop::Arg &aDP = a[na];
    WFD_ARG(DP, 3, true, false, a[na++]);
    float *t[6];
    for (int i = 0; i < 6; i++)
        if (!(t[i] = p->scratch(i))) return false;
    if (!dplus(p, aP.dev, t[0], t[1], t[2])) return false;
    if (!dplus(p, p->f[F_POLD].d, t[3], t[4], t[5])) return false;
    k_iso_acc_vb<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->f[F_V].d, p->f[F_B].d, aDP.dev, p->f[F_POLD].d,
                                                            F3{t[0], t[1], t[2]}, F3{t[3], t[4], t[5]},
                                                            gV ? a[0].dev : nullptr, agB.dev);
    WFD_LAUNCHED();
    if (!p->finish_args(a, na)) return false;
    return p->end(0);
}

static inline bool vti_acc_v(Prop *p, float *gV, float *DP, float *DM) {
    if (!p->begin(bit(F_V) | bit(F_B) | bit(F_POLD) | bit(F_MOLD))) return false;
    Prop::Arg a[3];
    WFD_ARG(gV, 0, true, true, a[0]);
    WFD_ARG(DP, 1, true, false, a[1]);
    WFD_ARG(DM, 2, true, false, a[2]);
    k_vti_acc_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->f[F_V].d, p->f[F_B].d, a[1].dev, a[2].dev,
                                                           p->f[F_POLD].d, p->f[F_MOLD].d, a[0].dev);
    WFD_LAUNCHED();
    if (!p->finish_args(a, 3)) return false;
    return p->end(0);
}

// VEA imaging (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:730-787, src/prop3DAcoTTIDenQ_DEO2_FDTD.h:733-790)
static inline bool vea_acc(Prop *p, float *gV, float *gE, float *gA, float *P, float *M, float *DP, float *DM) {
    if (!p->begin(p->model_mask() | bit(F_POLD) | bit(F_MOLD))) return false;
    Prop::Arg a[7];
    WFD_ARG(gV, 0, true, true, a[0]);
    WFD_ARG(gE, 1, true, true, a[1]);
    WFD_ARG(gA, 2, true, true, a[2]);
    WFD_ARG(P, 3, true, false, a[3]);
    WFD_ARG(M, 4, true, false, a[4]);
    WFD_ARG(DP, 5, true, false, a[5]);
    WFD_ARG(DM, 6, true, false, a[6]);
    float *t[12];
    for (int i = 0; i < 12; i++)
        if (!(t[i] = p->scratch(i))) return false;
    if (!rplus(p, a[3].dev, t[0], t[1], t[2])) return false;
    if (!rplus(p, a[4].dev, t[3], t[4], t[5])) return false;
    if (!rplus(p, p->f[F_POLD].d, t[6], t[7], t[8])) return false;
    if (!rplus(p, p->f[F_MOLD].d, t[9], t[10], t[11])) return false;
    k_vea_acc<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(
        p->n, p->f[F_V].d, p->f[F_B].d, p->f[F_ETA].d, p->f[F_F].d, a[5].dev, a[6].dev, p->f[F_POLD].d,
        p->f[F_MOLD].d, F3{t[0], t[1], t[2]}, F3{t[3], t[4], t[5]}, F3{t[6], t[7], t[8]}, F3{t[9], t[10], t[11]},
        a[0].dev, a[1].dev, a[2].dev);
    WFD_LAUNCHED();
    if (!p->finish_args(a, 7)) return false;
    return p->end(0);
}

// ---------------------------------------------------------------------------------------
// wavefield-separation imaging (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:673-765,
// src/prop3DAcoVTIDenQ_DEO2_FDTD.h:606-725): per (x,y) column, zero-pad to 2nz, transform
// with sign +1, zero half the spectrum, transform back with sign -1, g += 2b/v^3 Re(S R).
// The transforms are batched 1D C2C FFTs over a chunk of columns (hipFFT; FFTW's sign +1
// is hipFFT's BACKWARD direction, both unnormalised).
// ---------------------------------------------------------------------------------------

__global__ void k_sep_pack(long ncol, long nz, long nfft, long col0, float scale, const float *__restrict__ a,
                           float2 *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncol * nfft) return;
    const long c = i / nfft, j = i - c * nfft;
    out[i] = j < nz ? make_float2(scale * a[(col0 + c) * nz + j], 0.0f) : make_float2(0.0f, 0.0f);
}

// zero bins lo+1 .. lo+nfft/2-1 of every column
__global__ void k_sep_filter(long ncol, long nfft, long lo, float2 *__restrict__ s) {
    const long half = nfft / 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncol * (half - 1)) return;
    const long c = i / (half - 1), j = i - c * (half - 1) + 1;
    s[c * nfft + lo + j] = make_float2(0.0f, 0.0f);
}

__global__ void k_sep_accum(long ncol, long nz, long nfft, long col0, const float *__restrict__ v,
                            const float *__restrict__ b, const float2 *__restrict__ s1, const float2 *__restrict__ r1,
                            const float2 *__restrict__ s2, const float2 *__restrict__ r2, float *__restrict__ g) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncol * nz) return;
    const long c = i / nz, j = i - c * nz;
    const long k = (col0 + c) * nz + j;
    const float V = v[k], B = b[k];
    const float factor = 2 * B / (V * V * V);
    const float2 a = s1[c * nfft + j], r = r1[c * nfft + j];
    float re = a.x * r.x - a.y * r.y;
    if (s2) {
        const float2 a2 = s2[c * nfft + j], q2 = r2[c * nfft + j];
        re = re + (a2.x * q2.x - a2.y * q2.y);
    }
    g[k] += factor * re;
}

struct SepPlan {
    hipfftHandle plan = 0;
    long ncol = 0, nfft = 0;
    float2 *buf[4] = {};
    ~SepPlan() {
        if (plan) hipfftDestroy(plan);
        for (auto &b : buf)
            if (b) hipFree(b);
    }
};

static inline bool wavefield_sep(Prop *p, float *gV, float *DP, float *DM, long isFWI) {
    const bool two = p->phys != ISO;
    if (!p->begin(bit(F_V) | bit(F_B) | bit(F_POLD) | (two ? bit(F_MOLD) : 0u))) return false;
    Prop::Arg a[3];
    WFD_ARG(gV, 0, true, true, a[0]);
    WFD_ARG(DP, 1, true, false, a[1]);
    if (two) WFD_ARG(DM, 2, true, false, a[2]);
    const long nz = p->nz, nfft = 2 * nz, ncols = p->nx * p->ny;
    // chunk so each complex buffer stays near 128 MiB
    long chunk = std::max<long>(1, (128l << 20) / (long)(sizeof(float2) * nfft));
    chunk = std::min(chunk, ncols);
    SepPlan sp;
    int nfi = (int)nfft;
    if (hipfftPlanMany(&sp.plan, 1, &nfi, nullptr, 1, nfi, nullptr, 1, nfi, HIPFFT_C2C, (int)chunk) != HIPFFT_SUCCESS) {
        set_error("hipfftPlanMany failed");
        return false;
    }
    hipfftSetStream(sp.plan, p->stream);
    const int nbuf = two ? 4 : 2;
    for (int i = 0; i < nbuf; i++) WFD_TRY(hipMalloc(&sp.buf[i], sizeof(float2) * chunk * nfft));
    const float scale = 1.0f / (float)(nfft);
    const long kadj = isFWI ? 0 : nfft / 2;
    const float *src[4] = {a[1].dev, p->f[F_POLD].d, two ? a[2].dev : nullptr, two ? p->f[F_MOLD].d : nullptr};
    for (long col0 = 0; col0 < ncols; col0 += chunk) {
        const long nc = std::min(chunk, ncols - col0);
        for (int i = 0; i < nbuf; i++) {
            k_sep_pack<<<(unsigned)((nc * nfft + 255) / 256), 256, 0, p->stream>>>(nc, nz, nfft, col0, scale, src[i],
                                                                                   sp.buf[i]);
            WFD_LAUNCHED();
            if (hipfftExecC2C(sp.plan, (hipfftComplex *)sp.buf[i], (hipfftComplex *)sp.buf[i], HIPFFT_BACKWARD) !=
                HIPFFT_SUCCESS) {
                set_error("hipfftExecC2C failed");
                return false;
            }
            // nonlinear (even i) keeps bins 0..nz; adjoint (odd i) zeroes per FWI/RTM
            const long lo = (i % 2 == 0) ? nfft / 2 : kadj;
            if (nfft / 2 > 1) {
                k_sep_filter<<<(unsigned)((nc * (nfft / 2 - 1) + 255) / 256), 256, 0, p->stream>>>(nc, nfft, lo,
                                                                                                   sp.buf[i]);
                WFD_LAUNCHED();
            }
            if (hipfftExecC2C(sp.plan, (hipfftComplex *)sp.buf[i], (hipfftComplex *)sp.buf[i], HIPFFT_FORWARD) !=
                HIPFFT_SUCCESS) {
                set_error("hipfftExecC2C failed");
                return false;
            }
        }
        k_sep_accum<<<(unsigned)((nc * nz + 255) / 256), 256, 0, p->stream>>>(
            nc, nz, nfft, col0, p->f[F_V].d, p->f[F_B].d, sp.buf[0], sp.buf[1], two ? sp.buf[2] : nullptr,
            two ? sp.buf[3] : nullptr, a[0].dev);
        WFD_LAUNCHED();
    }
    WFD_TRY(hipStreamSynchronize(p->stream));
    if (!p->finish_args(a, two ? 3 : 2)) return false;
    return p->end(0);
}

// TimeStep with the Born loops' per-step call fused into it (additive entry points):
//   EPI_SAVE_DP   a0 = DP, a1 = DM (out)            TimeStep + ScaleSpatialDerivatives, into DP/DM
//   EPI_BORN_INJ  a0 = dV, a1 = DP, a2 = DM (in)    TimeStep + ForwardBornInjection_V
//   EPI_IMAGE     a0 = g (in/out), a1 = DP, a2 = DM TimeStep + AdjointBornAccumulation_V
// With the fused kernel switched off, the same results come from the two-pass step followed
// by the separate kernel.  pSpace/mSpace are not part of the result.
static inline bool born_step(Prop *p, int epi, float *a0, float *a1, float *a2) {
    const bool two = p->phys != ISO;
    if (!p->begin(p->all_mask())) return false;
    Prop::Arg a[3];
    int na = 0;
    if (epi == EPI_SAVE_DP) {
        WFD_ARG(a0, 0, false, true, a[na++]);
        if (two) WFD_ARG(a1, 1, false, true, a[na++]);
    } else {
        WFD_ARG(a0, 0, true, epi == EPI_IMAGE, a[na++]);
        WFD_ARG(a1, 1, true, false, a[na++]);
        if (two) WFD_ARG(a2, 2, true, false, a[na++]);
    }
    float *e0 = a[0].dev, *e1 = na > 1 ? a[1].dev : nullptr, *e2 = na > 2 ? a[2].dev : nullptr;
    if (p->fused) {
        bool ok = p->dim == 2             ? step_2d_epi(p, epi, e0, e1, e2)
                  : epi == EPI_SAVE_DP    ? step_fused_epi<EPI_SAVE_DP>(p, false, e0, e1, nullptr)
                  : epi == EPI_BORN_INJ ? step_fused_epi<EPI_BORN_INJ>(p, false, e0, e1, e2)
                                        : step_fused_epi<EPI_IMAGE>(p, false, e0, e1, e2);
        if (!ok) return false;
        p->swap_p();
    } else {
        if (!step_device(p, false)) return false;
        const Field *f = p->f;
        if (epi == EPI_SAVE_DP) {
            k_scale_space_into<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, f[F_V].d, f[F_B].d, f[F_PSPACE].d,
                                                                       two ? f[F_MSPACE].d : nullptr, e0, e1);
        } else if (epi == EPI_BORN_INJ) {
            if (two)
                k_vti_inj_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, f[F_V].d, e0, e1, e2, f[F_PCUR].d,
                                                                       f[F_MCUR].d);
            else if (p->dim == 2)  // the 2D text's V source: p += 2 dt^2 dV / v DP
                k2_iso_inj_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, f[F_V].d, e0, e1, f[F_PCUR].d);
            else
                k_iso_inj_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, f[F_V].d, f[F_B].d, e0, e1,
                                                                       f[F_PCUR].d);
        } else {
            if (two)
                k_vti_acc_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, f[F_V].d, f[F_B].d, e1, e2, f[F_POLD].d,
                                                                       f[F_MOLD].d, e0);
            else
                k_iso_acc_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, f[F_V].d, f[F_B].d, e1, f[F_POLD].d, e0);
        }
        WFD_LAUNCHED();
    }
    if (!p->finish_args(a, na)) return false;
    Mask w = bit(F_PCUR) | (two ? bit(F_MCUR) : 0);
    if (!p->fused) w |= bit(F_PSPACE) | (two ? bit(F_MSPACE) : 0);
    return p->end(w);
}

// Iso background and Born wavefields stepped together (additive): the same results as
// TimeStepSaveDP(DP) on the background (pCur/pOld) followed by TimeStepBornInjection_V(dV, DP)
// on a handle with the same model whose field is the Born field (mCur/mOld, allocated and
// zeroed at the first call; registered sources with fields bit 2 and receivers with fields 2
// then address it).  One pass reads the model once for both fields and forms DP in registers.
// Iso: the coupled Born step's second wavefield (the Born field) in the mCur / mOld slots,
// allocated and zeroed at its first use
static inline bool ensure_born_field(Prop *p) {
    if (!p->born_field) {
        for (int i : {F_MCUR, F_MOLD}) {
            WFD_TRY(hipMalloc(&p->f[i].d, sizeof(float) * p->n));
            WFD_TRY(hipMemsetAsync(p->f[i].d, 0, sizeof(float) * p->n, p->stream));
            p->f[i].used = true;
        }
        p->born_field = true;
    }
    return true;
}

static inline bool born_coupled(Prop *p, float *dV, float *DP) {
    if (p->phys != ISO) {
        set_error("the coupled Born step is defined for Iso only");
        return false;
    }
    if (!ensure_born_field(p)) return false;
    if (!p->begin(p->all_mask())) return false;
    Prop::Arg a[2];
    WFD_ARG(dV, 0, true, false, a[0]);
    WFD_ARG(DP, 1, false, true, a[1]);
    Field *f = p->f;
    if (p->fused) {
        // the new fields are in pOld and mOld: swap both pairs
        if (!step_fused_epi<EPI_BORN_COUPLED>(p, false, a[0].dev, a[1].dev, nullptr)) return false;
        p->swap_p();
        std::swap(f[F_MCUR], f[F_MOLD]);
    } else {
        // the two-pass step on each field in turn (the Born field through the p slots)
        if (!step_device(p, false)) return false;
        k_scale_space_into<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, f[F_V].d, f[F_B].d, f[F_PSPACE].d, nullptr,
                                                                   a[1].dev, nullptr);
        WFD_LAUNCHED();
        std::swap(f[F_PCUR], f[F_MCUR]);
        std::swap(f[F_POLD], f[F_MOLD]);
        const bool ok = step_device(p, false);
        if (ok) {
            k_iso_inj_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, f[F_V].d, f[F_B].d, a[0].dev, a[1].dev,
                                                                   f[F_PCUR].d);
        }
        std::swap(f[F_PCUR], f[F_MCUR]);  // each pair back in its slots, both already swapped
        std::swap(f[F_POLD], f[F_MOLD]);
        if (!ok) return false;
        WFD_LAUNCHED();
    }
    if (!p->finish_args(a, 2)) return false;
    return p->end(bit(F_PCUR) | (p->fused ? 0 : bit(F_PSPACE)));
}

// Several handles of one physics stepped together (additive): the same results as calling,
// on each handle i, TimeStep (epi 0) or the fused Born step epi (1 SAVE_DP, 2 BORN_INJ,
// 3 IMAGE) with the arrays a0[i], a1[i], a2[i] in born_step's order.  With the fused kernel
// on for every handle, one launch of k_fused_batch steps them all; it runs on the first
// handle's stream after the work queued on every handle's stream, and each handle's later
// work queues after it.  Handles must be distinct and in DEVICE coherence mode, the arrays
// device memory.
static inline bool time_step_batch(Physics phys, void **hs, long n, long epi, float **a0, float **a1,
                                   float **a2) {
    if (!hs || n < 1 || n > FUSED_MAX_BATCH) {
        set_error("TimeStepBatch: between 1 and " + std::to_string(FUSED_MAX_BATCH) + " handles");
        return false;
    }
    if (epi < EPI_NONE || epi > EPI_BORN_COUPLED || (epi == EPI_BORN_COUPLED && phys != ISO)) {
        set_error("TimeStepBatch: epilogue must be 0 (none), 1 (save DP), 2 (Born injection V), 3 (imaging V) or, "
                  "for Iso, 4 (the coupled Born step)");
        return false;
    }
    Prop *ps[FUSED_MAX_BATCH];
    bool all_fused = true;
    for (long i = 0; i < n; i++) {
        ps[i] = reinterpret_cast<Prop *>(hs[i]);
        if (!ps[i] || ps[i]->phys != phys) {
            set_error("TimeStepBatch: null handle or a handle of another physics");
            return false;
        }
        for (long j = 0; j < i; j++)
            if (ps[j] == ps[i]) {
                set_error("TimeStepBatch: a handle appears twice");
                return false;
            }
        if (ps[i]->coherence != DEVICE) {
            set_error("TimeStepBatch: handles must be in DEVICE coherence mode");
            return false;
        }
        all_fused = all_fused && ps[i]->fused;
    }
    const bool two = phys != ISO;
    // arrays the epilogue needs: SAVE_DP (DP[, DM]) = a0[, a1]; coupled (dV, DP) = a0, a1; others
    // (x, DP[, DM]) = a0, a1[, a2]
    const int need = epi == EPI_NONE ? 0 : epi == EPI_BORN_COUPLED ? 2 : (epi == EPI_SAVE_DP ? 1 : 2) + (two ? 1 : 0);
    float **arr[3] = {a0, a1, a2};
    for (int k = 0; k < need; k++) {
        if (!arr[k]) {
            set_error("TimeStepBatch: missing array list");
            return false;
        }
        for (long i = 0; i < n; i++)
            if (!arr[k][i] || !is_device_ptr(arr[k][i])) {
                set_error("TimeStepBatch: epilogue arrays must be device memory");
                return false;
            }
    }
    auto get = [&](int k, long i) -> float * { return k < need ? arr[k][i] : nullptr; };
    if (!all_fused) {  // the same results, one handle after the other
        for (long i = 0; i < n; i++) {
            const bool ok = epi == EPI_NONE           ? time_step(ps[i], false)
                            : epi == EPI_BORN_COUPLED ? born_coupled(ps[i], get(0, i), get(1, i))
                                                      : born_step(ps[i], (int)epi, get(0, i), get(1, i), get(2, i));
            if (!ok) return false;
        }
        return true;
    }
    if (epi == EPI_BORN_COUPLED)
        for (long i = 0; i < n; i++)
            if (!ensure_born_field(ps[i])) return false;
    hipStream_t s0 = ps[0]->stream;
    for (long i = 0; i < n; i++)
        if (!ps[i]->ev) WFD_TRY(hipEventCreateWithFlags(&ps[i]->ev, hipEventDisableTiming));
    for (long i = 1; i < n; i++)
        if (ps[i]->stream != s0) {
            WFD_TRY(hipEventRecord(ps[i]->ev, ps[i]->stream));
            WFD_TRY(hipStreamWaitEvent(s0, ps[i]->ev, 0));
        }
    float *e[3][FUSED_MAX_BATCH] = {};
    for (int k = 0; k < need; k++)
        for (long i = 0; i < n; i++) e[k][i] = arr[k][i];
    const int nn = (int)n;
    const bool ok = epi == EPI_NONE        ? step_fused_batch<EPI_NONE>(ps, nn, nullptr, nullptr, nullptr)
                    : epi == EPI_SAVE_DP   ? step_fused_batch<EPI_SAVE_DP>(ps, nn, e[0], e[1], nullptr)
                    : epi == EPI_BORN_INJ  ? step_fused_batch<EPI_BORN_INJ>(ps, nn, e[0], e[1], e[2])
                    : epi == EPI_IMAGE     ? step_fused_batch<EPI_IMAGE>(ps, nn, e[0], e[1], e[2])
                                           : step_fused_batch<EPI_BORN_COUPLED>(ps, nn, e[0], e[1], nullptr);
    if (!ok) return false;
    WFD_TRY(hipEventRecord(ps[0]->ev, s0));
    for (long i = 1; i < n; i++)
        if (ps[i]->stream != s0) WFD_TRY(hipStreamWaitEvent(ps[i]->stream, ps[0]->ev, 0));
    for (long i = 0; i < n; i++) {
        ps[i]->swap_p();
        if (epi == EPI_BORN_COUPLED) std::swap(ps[i]->f[F_MCUR], ps[i]->f[F_MOLD]);  // the Born field's pair
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// 2D propagators: Born sources and imaging (src/prop2DAcoIsoDenQ_DEO2_FDTD.h:249-582)
// ---------------------------------------------------------------------------------------

static inline bool iso2_inj_v(Prop *p, float *dV, float *DP) {
    if (!p->begin(bit(F_V) | bit(F_PCUR))) return false;
    Prop::Arg a[2];
    WFD_ARG(dV, 0, true, false, a[0]);
    WFD_ARG(DP, 1, true, false, a[1]);
    k2_iso_inj_v<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, p->f[F_V].d, a[0].dev, a[1].dev,
                                                            p->f[F_PCUR].d);
    WFD_LAUNCHED();
    if (!p->finish_args(a, 2)) return false;
    return p->end(bit(F_PCUR));
}

// VB (dV != null) and B (dV == null): l = D-(dB D+ P)
static inline bool iso2_inj_vb(Prop *p, float *dV, float *dB, float *P, float *DP) {
    if (!p->begin(bit(F_V) | bit(F_B) | bit(F_PCUR))) return false;
    Prop::Arg a[4];
    int na = 0;
    if (dV) WFD_ARG(dV, 0, true, false, a[na++]);
    Prop::Arg &adB = a[na];
    WFD_ARG(dB, 1, true, false, a[na++]);
    Prop::Arg &aP = a[na];
    WFD_ARG(P, 2, true, false, a[na++]);
    Prop::Arg &aDP = a[na];
    WFD_ARG(DP, 3, true, false, a[na++]);
    float *t[4];
    for (int i = 0; i < 4; i++)
        if (!(t[i] = p->scratch(i))) return false;
    if (!dplus_2d(p, aP.dev, t[0], t[1])) return false;
    k2_scale2<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, adB.dev, t[0], t[1]);
    WFD_LAUNCHED();
    if (!dminus_2d(p, t[0], t[1], t[2], t[3])) return false;
    k2_iso_inj_vb<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, p->f[F_V].d, p->f[F_B].d,
                                                             dV ? a[0].dev : nullptr, adB.dev, aDP.dev, t[2], t[3],
                                                             p->f[F_PCUR].d);
    WFD_LAUNCHED();
    if (!p->finish_args(a, na)) return false;
    return p->end(bit(F_PCUR));
}

// imaging V (gB null), VB, B (gV null)
static inline bool iso2_acc(Prop *p, float *gV, float *gB, float *P, float *DP) {
    if (!p->begin(bit(F_V) | bit(F_B) | bit(F_POLD))) return false;
    Prop::Arg a[4];
    int na = 0;
    Prop::Arg *agV = nullptr, *agB = nullptr, *aP = nullptr;
    if (gV) {
        agV = &a[na];
        WFD_ARG(gV, 0, true, true, a[na++]);
    }
    if (gB) {
        agB = &a[na];
        WFD_ARG(gB, 1, true, true, a[na++]);
        aP = &a[na];
        WFD_ARG(P, 2, true, false, a[na++]);
    }
    Prop::Arg &aDP = a[na];
    WFD_ARG(DP, 3, true, false, a[na++]);
    float *t[4] = {};
    if (gB) {
        for (int i = 0; i < 4; i++)
            if (!(t[i] = p->scratch(i))) return false;
        if (!dplus_2d(p, aP->dev, t[0], t[1])) return false;
        if (!dplus_2d(p, p->f[F_POLD].d, t[2], t[3])) return false;
    }
    k2_iso_acc<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->f[F_V].d, p->f[F_B].d, aDP.dev, p->f[F_POLD].d,
                                                          t[0], t[1], t[2], t[3], agV ? agV->dev : nullptr,
                                                          agB ? agB->dev : nullptr);
    WFD_LAUNCHED();
    if (!p->finish_args(a, na)) return false;
    return p->end(0);
}

// 2D VTI/TTI VEA source (src/prop2DAcoVTIDenQ_DEO2_FDTD.h:383-462,
// src/prop2DAcoTTIDenQ_DEO2_FDTD.h:362-441): D+ (rotated for TTI) of P and M, the perturbation
// sandwich, D- (rotated for TTI), then p += dt^2 v^2/b (2b dV/v^3 DP + lx + lz), m likewise
static inline bool aniso2_inj_vea(Prop *p, float *dV, float *dE, float *dA, float *P, float *M, float *DP,
                                  float *DM) {
    if (!p->begin(p->model_mask() | bit(F_PCUR) | bit(F_MCUR))) return false;
    Prop::Arg a[7];
    WFD_ARG(dV, 0, true, false, a[0]);
    WFD_ARG(dE, 1, true, false, a[1]);
    WFD_ARG(dA, 2, true, false, a[2]);
    WFD_ARG(P, 3, true, false, a[3]);
    WFD_ARG(M, 4, true, false, a[4]);
    WFD_ARG(DP, 5, true, false, a[5]);
    WFD_ARG(DM, 6, true, false, a[6]);
    float *t[8];
    for (int i = 0; i < 8; i++)
        if (!(t[i] = p->scratch(i))) return false;
    if (!aniso2_dplus(p, a[3].dev, t[0], t[1])) return false;
    if (!aniso2_dplus(p, a[4].dev, t[2], t[3])) return false;
    k2_vea_inj_mid<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->f[F_B].d, p->f[F_ETA].d, p->f[F_F].d,
                                                              a[1].dev, a[2].dev, t[0], t[1], t[3], t[4], t[5], t[6],
                                                              t[7]);
    WFD_LAUNCHED();
    if (!aniso2_dminus(p, t[4], t[5], t[0], t[1])) return false;
    if (!aniso2_dminus(p, t[6], t[7], t[2], t[3])) return false;
    k2_vea_inj_final<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->dt, p->f[F_V].d, p->f[F_B].d, a[0].dev,
                                                                a[5].dev, a[6].dev, t[0], t[1], t[2], t[3],
                                                                p->f[F_PCUR].d, p->f[F_MCUR].d);
    WFD_LAUNCHED();
    if (!p->finish_args(a, 7)) return false;
    return p->end(bit(F_PCUR) | bit(F_MCUR));
}

// 2D VTI/TTI VEA imaging (src/prop2DAcoVTIDenQ_DEO2_FDTD.h:629-686,
// src/prop2DAcoTTIDenQ_DEO2_FDTD.h:607-666)
static inline bool aniso2_acc_vea(Prop *p, float *gV, float *gE, float *gA, float *P, float *M, float *DP,
                                  float *DM) {
    if (!p->begin(p->model_mask() | bit(F_POLD) | bit(F_MOLD))) return false;
    Prop::Arg a[7];
    WFD_ARG(gV, 0, true, true, a[0]);
    WFD_ARG(gE, 1, true, true, a[1]);
    WFD_ARG(gA, 2, true, true, a[2]);
    WFD_ARG(P, 3, true, false, a[3]);
    WFD_ARG(M, 4, true, false, a[4]);
    WFD_ARG(DP, 5, true, false, a[5]);
    WFD_ARG(DM, 6, true, false, a[6]);
    float *t[8];
    for (int i = 0; i < 8; i++)
        if (!(t[i] = p->scratch(i))) return false;
    if (!aniso2_dplus(p, a[3].dev, t[0], t[1])) return false;
    if (!aniso2_dplus(p, a[4].dev, t[2], t[3])) return false;
    if (!aniso2_dplus(p, p->f[F_POLD].d, t[4], t[5])) return false;
    if (!aniso2_dplus(p, p->f[F_MOLD].d, t[6], t[7])) return false;
    k2_vea_acc<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(
        p->n, p->f[F_V].d, p->f[F_B].d, p->f[F_ETA].d, p->f[F_F].d, a[5].dev, a[6].dev, p->f[F_POLD].d,
        p->f[F_MOLD].d, t[0], t[1], t[3], t[4], t[5], t[7], a[0].dev, a[1].dev, a[2].dev);
    WFD_LAUNCHED();
    if (!p->finish_args(a, 7)) return false;
    return p->end(0);
}

// Prop2DAcoVTIDenQ_DEO2_FDTD_applyFirstDerivatives2D_PlusHalf: the stateless D+ on host arrays
// with the caller's coefficients (src/prop2DAcoVTIDenQ_DEO2_FDTD.cc:154-175); staged through
// device buffers on a stream of its own, synchronous like the reference
static inline bool dplus_2d_host(long fs, long nx, long nz, float c1, float c2, float c3, float c4, float invDx,
                                 float invDz, const float *inX, const float *inZ, float *outX, float *outZ) {
    if (nx <= 0 || nz <= 0 || !inX || !inZ || !outX || !outZ) {
        set_error("applyFirstDerivatives2D_PlusHalf: empty grid or null array");
        return false;
    }
    const size_t bytes = sizeof(float) * (size_t)nx * (size_t)nz;
    Grid g{};
    g.nx = nx;
    g.ny = 1;
    g.nz = nz;
    g.nynz = nz;
    g.fs = fs ? 1 : 0;
    g.idx = invDx;
    g.idy = invDx;
    g.idz = invDz;
    hipStream_t st = nullptr;
    float *d[4] = {};
    bool ok = false;
    do {
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) break;
        bool alloc = true;
        for (auto &x : d) alloc = alloc && hipMalloc(&x, bytes) == hipSuccess;
        if (!alloc) break;
        if (hipMemcpyAsync(d[0], inX, bytes, hipMemcpyHostToDevice, st) != hipSuccess) break;
        if (hipMemcpyAsync(d[1], inZ, bytes, hipMemcpyHostToDevice, st) != hipSuccess) break;
        k2_dplus_coef<<<op2_grid(g), op2_block(), 0, st>>>(g, make_float4(c1, c2, c3, c4), d[0], d[1], d[2], d[3]);
        if (hipGetLastError() != hipSuccess) break;
        if (hipMemcpyAsync(outX, d[2], bytes, hipMemcpyDeviceToHost, st) != hipSuccess) break;
        if (hipMemcpyAsync(outZ, d[3], bytes, hipMemcpyDeviceToHost, st) != hipSuccess) break;
        ok = hipStreamSynchronize(st) == hipSuccess;
    } while (false);
    if (!ok) set_error(std::string("applyFirstDerivatives2D_PlusHalf: ") + hipGetErrorString(hipGetLastError()));
    for (auto &x : d)
        if (x) (void)hipFree(x);
    if (st) (void)hipStreamDestroy(st);
    return ok;
}

}  // namespace wfd

