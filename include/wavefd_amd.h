/*
 * wavefd_amd.h -- additive entry points of the three MI355X propagator libraries.  None has
 * a counterpart in the reference; each library exports them with its own prefix
 * (Prop3DAcoIsoDenQ_DEO2_FDTD_, Prop3DAcoVTIDenQ_DEO2_FDTD_, Prop3DAcoTTIDenQ_DEO2_FDTD_).
 * They serve callers that keep the time loop on the GPU: device pointers, stream-ordered
 * calls, registered sources/receivers and one-call step+inject+extract.
 *
 * Host coherence.  The reference hands out host arrays (getters) that callers read and
 * write between calls.  Each handle has a mode:
 *   1 = COHERENT (default): before a call every host array already handed out by a getter
 *       and read by the call is copied to the device; after it, every such array the call
 *       writes is copied back, and the call returns when the work is done.
 *   0 = DEVICE: no implicit copies and calls are stream-ordered (they may return before the
 *       work ends); use SyncToDevice / SyncToHost / Synchronize and GetDevicePtr.
 * Caller-owned arrays passed to Born / imaging calls may be host or device memory in either
 * mode; host arrays are staged through device scratch.
 */
#ifndef WAVEFD_AMD_H
#define WAVEFD_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

struct SourcePoint32; /* spacetime.h */

#define WAVEFD_AMD_DECLARE(P)                                                                       \
    /* set / get the host-coherence mode of a handle (1 coherent, 0 device) */                     \
    void P##_SetHostCoherence(void *p, long mode);                                                  \
    long P##_GetHostCoherence(void *p);                                                             \
    /* copy every host array handed out so far to the device / from the device, and wait */        \
    void P##_SyncToDevice(void *p);                                                                 \
    void P##_SyncToHost(void *p);                                                                   \
    /* wait for all work queued on the handle's stream */                                          \
    void P##_Synchronize(void *p);                                                                  \
    /* device address of a state array by name: v b dtOmegaInvQ eps eta f sinTheta cosTheta       \
     * sinPhi cosPhi pCur pOld mCur mOld pSpace mSpace; NULL if the physics has no such array */  \
    float *P##_GetDevicePtr(void *p, const char *name);                                             \
    /* the handle's HIP stream (hipStream_t) */                                                     \
    void *P##_GetStream(void *p);                                                                   \
    /* register injection points (as partitions, e.g. from source_blocking) and their data       \
     * (Julia (nt, nr), host or device); fields: 1 = P, 2 = M, 3 = both.  0 on success. */       \
    long P##_SetSources(void *p, struct SourcePoint32 **partitions, size_t npartitions, long *npoints, \
                        float *data, size_t nt, long fields);                                       \
    /* register extraction points; a zeroed (nt, nr) trace buffer is kept on the device;         \
     * fields: 1 = P, 2 = M.  0 on success. */                                                     \
    long P##_SetReceivers(void *p, struct SourcePoint32 **partitions, size_t npartitions, long *npoints, \
                          size_t nt, long fields);                                                  \
    /* the same with the time interpolation fused in: the source data (Julia (nd, nr)) is at the  \
     * coarse rate, and each injection at fine step it adds the reference's interpadjoint of it   \
     * (src/spacetime.cc:4-62); the receivers' trace buffer (nd, nr) is at the coarse rate, and   \
     * each extraction at fine step it adds the interpforward contributions of that fine sample  \
     * (src/spacetime.cc:64-127).  h: nh x 8 filters (interpfilters, src/spacetime.jl:333-364), \
     * host or device; nd >= 7 coarse samples, nm fine samples.  The results equal the host       \
     * interpolation followed by the fine-rate calls, bit for bit.  0 on success. */             \
    long P##_SetSourcesInterp(void *p, struct SourcePoint32 **partitions, size_t npartitions, long *npoints, \
                              float *data, size_t nd, float *h, size_t nh, size_t nm, long fields);  \
    long P##_SetReceiversInterp(void *p, struct SourcePoint32 **partitions, size_t npartitions, long *npoints, \
                                size_t nd, float *h, size_t nh, size_t nm, long fields);            \
    /* TimeStep, then inject the sources at it (1-based) into the new pCur (mCur), then extract   \
     * the receivers at it from the new pCur: the loop body of test/spacetime.jl:486-492 */       \
    void P##_TimeStepInjectExtract(void *p, long it);                                               \
    /* nsteps of TimeStepInjectExtract (it0 .. it0+nsteps-1) as a HIP graph of two steps, captured  \
     * once on the handle's stream (the time index in a device counter the graph advances) and      \
     * launched nsteps/2 times; the same results as the calls one by one.  DEVICE mode and the      \
     * fused step only; time indices checked first; returns when the steps are done.  0 on          \
     * success. */                                                                                  \
    long P##_TimeStepInjectExtractGraph(void *p, long it0, long nsteps);                            \
    /* the two halves alone, for loops that put other work between the step and the             \
     * extraction (Born: step, Born source, extract; adjoint: step, adjoint source, image).       \
     * 0 on success. */                                                                             \
    long P##_InjectRegistered(void *p, long it);                                                    \
    long P##_ExtractRegistered(void *p, long it);                                                   \
    /* copy the receiver traces (Julia (nt, nr)) to out (host or device).  0 on success. */       \
    long P##_GetReceiverData(void *p, float *out);                                                  \
    /* 1 (default): fused single-pass step kernel; 0: two-pass operator kernels */                 \
    void P##_SetFusedStep(void *p, long fused);                                                     \
    /* tile shape of the fused step kernel: 0 automatic, 1 = 64 (z) x 16 (y) threads, 2 = 32 x 32;  \
     * results do not depend on it.  0 on success, -1 for another value. */                         \
    long P##_SetFusedTile(void *p, long tile);                                                      \
    long P##_GetFusedTile(void *p);                                                                 \
    /* the fused step's whole launch configuration: tile as above, rounds in step (lockstep)       \
     * -1 automatic / 0 off / 1 on (only where the grid has 16 rounds of workgroups), x-planes per  \
     * chunk (0 automatic); results do not depend on it.  0 on success.  GetFusedConfig writes     \
     * what a plain TimeStep launches with: out[0] tile, out[1] lockstep, out[2] chunk,             \
     * out[3] workgroups.  The autotuner (wavefd_amd/tuning.py) sets it.  */                      \
    long P##_SetFusedConfig(void *p, long tile, long lockstep, long chunk);                         \
    long P##_GetFusedConfig(void *p, long *out);                                                    \
    /* step n (1..8) distinct handles of this physics, all in DEVICE mode, by one launch: the     \
     * same results as calling on each handle i TimeStep (epilogue 0), TimeStepSaveDP (1: a0[i] =  \
     * DP, a1[i] = DM), TimeStepBornInjection_V (2: a0 = dV, a1 = DP, a2 = DM) or                  \
     * TimeStepAdjointBornAccumulation_V (3: a0 = g, a1 = DP, a2 = DM) or, Iso only,             \
     * TimeStepBornCoupled_V (4: a0 = dV, a1 = DP); arrays device memory,                         \
     * DM lists only for VTI/TTI.  Runs after the work queued on every handle's stream; later     \
     * work on each handle's stream runs after it.  0 on success. */                              \
    long P##_TimeStepBatch(void **ps, long n, long epilogue, float **a0, float **a1, float **a2);  \
    /* last error message of the library ("" if none) and reset */                                 \
    const char *P##_LastError(void);                                                                \
    void P##_ClearError(void);

WAVEFD_AMD_DECLARE(Prop3DAcoIsoDenQ_DEO2_FDTD)
WAVEFD_AMD_DECLARE(Prop3DAcoVTIDenQ_DEO2_FDTD)
WAVEFD_AMD_DECLARE(Prop3DAcoTTIDenQ_DEO2_FDTD)

#undef WAVEFD_AMD_DECLARE

/* Born loops' per-step pairs as one fused step; no counterpart in the reference (its callers
 * make the two calls).  Same results as the pairs, on the same arguments:
 *   TimeStepSaveDP(p, DP[, DM])                  = TimeStep; ScaleSpatialDerivatives, with the
 *                                                  scaled pSpace (mSpace) written to DP (DM)
 *   TimeStepBornInjection_V(p, dV, DP[, DM])     = TimeStep; ForwardBornInjection_V(dV, DP[, DM])
 *   TimeStepAdjointBornAccumulation_V(p, g, DP[, DM])
 *                                                = TimeStep; AdjointBornAccumulation_V(g, DP[, DM])
 * pSpace/mSpace are not part of the result (the fused kernel leaves them as they were).
 * Arrays: host or device, like the reference's Born calls.  0 on success. */
long Prop3DAcoIsoDenQ_DEO2_FDTD_TimeStepSaveDP(void *p, float *DP);
long Prop3DAcoIsoDenQ_DEO2_FDTD_TimeStepBornInjection_V(void *p, float *dVel, float *DP);
long Prop3DAcoIsoDenQ_DEO2_FDTD_TimeStepAdjointBornAccumulation_V(void *p, float *dVel, float *DP);
/* Iso background and Born wavefields in one pass; no counterpart in the reference.  The same
 * results as TimeStepSaveDP(p, DP) on the background field (pCur/pOld) followed by
 * TimeStepBornInjection_V(q, dVel, DP) on a handle q with the same model holding the Born field,
 * which this handle keeps as mCur/mOld (allocated and zeroed at the first call; GetDevicePtr
 * "mCur"/"mOld" return it, registered sources with fields bit 2 inject into it and registered
 * receivers with fields 2 extract from it).  0 on success. */
long Prop3DAcoIsoDenQ_DEO2_FDTD_TimeStepBornCoupled_V(void *p, float *dVel, float *DP);
long Prop3DAcoVTIDenQ_DEO2_FDTD_TimeStepSaveDP(void *p, float *DP, float *DM);
long Prop3DAcoVTIDenQ_DEO2_FDTD_TimeStepBornInjection_V(void *p, float *dVel, float *DP, float *DM);
long Prop3DAcoVTIDenQ_DEO2_FDTD_TimeStepAdjointBornAccumulation_V(void *p, float *dVel, float *DP, float *DM);
long Prop3DAcoTTIDenQ_DEO2_FDTD_TimeStepSaveDP(void *p, float *DP, float *DM);
long Prop3DAcoTTIDenQ_DEO2_FDTD_TimeStepBornInjection_V(void *p, float *dVel, float *DP, float *DM);
long Prop3DAcoTTIDenQ_DEO2_FDTD_TimeStepAdjointBornAccumulation_V(void *p, float *dVel, float *DP, float *DM);

#ifdef __cplusplus
}
#endif
#endif
