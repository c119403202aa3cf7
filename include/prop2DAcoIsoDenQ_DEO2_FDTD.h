/* libprop2DAcoIsoDenQ_DEO2_FDTD -- MI355X drop-in for the reference's 2D Iso library
 * (src/prop2DAcoIsoDenQ_DEO2_FDTD.cc, CMake target src/CMakeLists.txt).  Same names, argument
 * order and C types; arrays (nz, nx) float32, Julia column-major, k = kx nz + kz.  Calls are
 * synchronous and getters return host arrays kept coherent with the device (the additive
 * coherence switch at the end makes calls stream-ordered instead). */
#ifndef PROP2DACOISODENQ_DEO2_FDTD_AMD_H
#define PROP2DACOISODENQ_DEO2_FDTD_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* src/prop2DAcoIsoDenQ_DEO2_FDTD.cc:6-23 */
void *Prop2DAcoIsoDenQ_DEO2_FDTD_alloc(long fs, long nthread, long nx, long nz, long nsponge, float dx, float dz,
                                       float dt, long nbx, long nbz);
/* :25-28 */
void Prop2DAcoIsoDenQ_DEO2_FDTD_free(void *p);
/* :30-39 (src/absorb.h:23-75; exits the process on freqQ < FLT_EPSILON, like the reference) */
void Prop2DAcoIsoDenQ_DEO2_FDTD_SetupDtOmegaInvQ(void *p, float freqQ, float qMin, float qInterior);
/* :41-44 */
void Prop2DAcoIsoDenQ_DEO2_FDTD_TimeStep(void *p);
/* :46-49 */
void Prop2DAcoIsoDenQ_DEO2_FDTD_ScaleSpatialDerivatives(void *p);
/* :51-55, :57-61, :63-67 */
void Prop2DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_V(void *p, float *dVel, float *wavefieldDP);
void Prop2DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_VB(void *p, float *dVel, float *dBuoy, float *wavefieldP,
                                                        float *wavefieldDP);
void Prop2DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_B(void *p, float *dBuoy, float *wavefieldP, float *wavefieldDP);
/* :69-73, :75-79, :81-85 */
void Prop2DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_V(void *p, float *dVel, float *wavefieldDP);
void Prop2DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_VB(void *p, float *dVel, float *dBuoy, float *wavefieldP,
                                                           float *wavefieldDP);
void Prop2DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_B(void *p, float *dBuoy, float *wavefieldP,
                                                          float *wavefieldDP);
/* :87-91 */
void Prop2DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_wavefieldsep(void *p, float *dVel, float *wavefieldDP,
                                                                     const long isFWI);
/* :93-101 */
long Prop2DAcoIsoDenQ_DEO2_FDTD_getNx(void *p);
long Prop2DAcoIsoDenQ_DEO2_FDTD_getNz(void *p);
/* :103-131 */
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getV(void *p);
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getB(void *p);
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getDtOmegaInvQ(void *p);
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getPSpace(void *p);
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getPCur(void *p);
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getPOld(void *p);

/* additive, no counterpart in the reference: host coherence (1 coherent, 0 device), explicit
 * copies, device addresses by name (v b dtOmegaInvQ pCur pOld pSpace), errors */
void Prop2DAcoIsoDenQ_DEO2_FDTD_SetHostCoherence(void *p, long mode);
void Prop2DAcoIsoDenQ_DEO2_FDTD_SyncToDevice(void *p);
void Prop2DAcoIsoDenQ_DEO2_FDTD_SyncToHost(void *p);
float *Prop2DAcoIsoDenQ_DEO2_FDTD_GetDevicePtr(void *p, const char *name);
/* 1 (default): the fused single-pass step kernel; 0: the two passes (sandwich, then D- and
 * the update); the results are the same */
void Prop2DAcoIsoDenQ_DEO2_FDTD_SetFusedStep(void *p, long fused);
/* the Born loops' per-step pairs as one step; no counterpart in the reference (its callers
 * make the two calls).  Same results as the pairs, on the same arguments:
 *   TimeStepSaveDP(p, DP)                   = TimeStep; ScaleSpatialDerivatives, with the
 *                                                 scaled pSpace written to DP
 *   TimeStepBornInjection_V(p, dV, DP)      = TimeStep; ForwardBornInjection_V(dV, DP)
 *   TimeStepAdjointBornAccumulation_V(p, g, DP)
 *                                               = TimeStep; AdjointBornAccumulation_V(g, DP)
 * pSpace is not part of the result.  Arrays: host or device.  0 on success. */
long Prop2DAcoIsoDenQ_DEO2_FDTD_TimeStepSaveDP(void *p, float *DP);
long Prop2DAcoIsoDenQ_DEO2_FDTD_TimeStepBornInjection_V(void *p, float *dVel, float *DP);
long Prop2DAcoIsoDenQ_DEO2_FDTD_TimeStepAdjointBornAccumulation_V(void *p, float *dVel, float *DP);
/* the handle's HIP stream (device-mode calls are ordered on it) and a wait for it */
void *Prop2DAcoIsoDenQ_DEO2_FDTD_GetStream(void *p);
void Prop2DAcoIsoDenQ_DEO2_FDTD_Synchronize(void *p);
const char *Prop2DAcoIsoDenQ_DEO2_FDTD_LastError(void);
void Prop2DAcoIsoDenQ_DEO2_FDTD_ClearError(void);

#ifdef __cplusplus
}
#endif
#endif
