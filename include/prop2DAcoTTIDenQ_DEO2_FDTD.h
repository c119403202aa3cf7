/* libprop2DAcoTTIDenQ_DEO2_FDTD -- MI355X drop-in for the reference's 2D TTI library
 * (src/prop2DAcoTTIDenQ_DEO2_FDTD.cc).  Same names, argument order and C types; arrays
 * (nz, nx) float32, Julia column-major, k = kx nz + kz.  The tilt is given as sinTheta and
 * cosTheta arrays.  Calls are synchronous and getters return host arrays kept coherent with
 * the device (the additive coherence switch at the end makes calls stream-ordered instead). */
#ifndef PROP2DACOTTIDENQ_DEO2_FDTD_AMD_H
#define PROP2DACOTTIDENQ_DEO2_FDTD_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* src/prop2DAcoTTIDenQ_DEO2_FDTD.cc:6-23 */
void *Prop2DAcoTTIDenQ_DEO2_FDTD_alloc(long fs, long nthread, long nx, long nz, long nsponge, float dx, float dz,
                                       float dt, long nbx, long nbz);
/* :25-28 */
void Prop2DAcoTTIDenQ_DEO2_FDTD_free(void *p);
/* :30-39 (src/absorb.h:23-75; exits the process on freqQ < FLT_EPSILON, like the reference) */
void Prop2DAcoTTIDenQ_DEO2_FDTD_SetupDtOmegaInvQ(void *p, float freqQ, float qMin, float qInterior);
/* :41-44 (the 2D TTI library has no linear step) */
void Prop2DAcoTTIDenQ_DEO2_FDTD_TimeStep(void *p);
/* :46-49 */
void Prop2DAcoTTIDenQ_DEO2_FDTD_ScaleSpatialDerivatives(void *p);
/* :51-55, :57-62 */
void Prop2DAcoTTIDenQ_DEO2_FDTD_ForwardBornInjection_V(void *p, float *dVel, float *wavefieldDP, float *wavefieldDM);
void Prop2DAcoTTIDenQ_DEO2_FDTD_ForwardBornInjection_VEA(void *p, float *dVel, float *dEps, float *dEta,
                                                         float *wavefieldP, float *wavefieldM, float *wavefieldDP,
                                                         float *wavefieldDM);
/* :64-68, :70-74 (hipFFT), :76-81 */
void Prop2DAcoTTIDenQ_DEO2_FDTD_AdjointBornAccumulation_V(void *p, float *dVel, float *wavefieldDP,
                                                          float *wavefieldDM);
void Prop2DAcoTTIDenQ_DEO2_FDTD_AdjointBornAccumulation_wavefieldsep_V(void *p, float *dVel, float *wavefieldDP,
                                                                       float *wavefieldDM, const long isFWI);
void Prop2DAcoTTIDenQ_DEO2_FDTD_AdjointBornAccumulation_VEA(void *p, float *dVel, float *dEps, float *dEta,
                                                            float *wavefieldP, float *wavefieldM, float *wavefieldDP,
                                                            float *wavefieldDM);
/* :83-91 */
long Prop2DAcoTTIDenQ_DEO2_FDTD_getNx(void *p);
long Prop2DAcoTTIDenQ_DEO2_FDTD_getNz(void *p);
/* :93-161 */
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getV(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getEps(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getEta(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getSinTheta(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getCosTheta(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getB(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getF(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getDtOmegaInvQ(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getPSpace(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getMSpace(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getPCur(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getPOld(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getMCur(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_getMOld(void *p);

/* additive, no counterpart in the reference: host coherence (1 coherent, 0 device), explicit
 * copies, device addresses by name (v b dtOmegaInvQ eps eta f sinTheta cosTheta pCur pOld mCur
 * mOld pSpace mSpace), errors */
void Prop2DAcoTTIDenQ_DEO2_FDTD_SetHostCoherence(void *p, long mode);
void Prop2DAcoTTIDenQ_DEO2_FDTD_SyncToDevice(void *p);
void Prop2DAcoTTIDenQ_DEO2_FDTD_SyncToHost(void *p);
float *Prop2DAcoTTIDenQ_DEO2_FDTD_GetDevicePtr(void *p, const char *name);
/* 1 (default): the fused single-pass step kernel; 0: the two passes (sandwich, then D- and
 * the update); the results are the same */
void Prop2DAcoTTIDenQ_DEO2_FDTD_SetFusedStep(void *p, long fused);
/* the Born loops' per-step pairs as one step; no counterpart in the reference (its callers
 * make the two calls).  Same results as the pairs, on the same arguments:
 *   TimeStepSaveDP(p, DP, DM)                   = TimeStep; ScaleSpatialDerivatives, with the
 *                                                 scaled pSpace (mSpace) written to DP (DM)
 *   TimeStepBornInjection_V(p, dV, DP, DM)      = TimeStep; ForwardBornInjection_V(dV, DP, DM)
 *   TimeStepAdjointBornAccumulation_V(p, g, DP, DM)
 *                                               = TimeStep; AdjointBornAccumulation_V(g, DP, DM)
 * pSpace/mSpace are not part of the result.  Arrays: host or device.  0 on success. */
long Prop2DAcoTTIDenQ_DEO2_FDTD_TimeStepSaveDP(void *p, float *DP, float *DM);
long Prop2DAcoTTIDenQ_DEO2_FDTD_TimeStepBornInjection_V(void *p, float *dVel, float *DP, float *DM);
long Prop2DAcoTTIDenQ_DEO2_FDTD_TimeStepAdjointBornAccumulation_V(void *p, float *dVel, float *DP, float *DM);
/* the handle's HIP stream (device-mode calls are ordered on it) and a wait for it */
void *Prop2DAcoTTIDenQ_DEO2_FDTD_GetStream(void *p);
void Prop2DAcoTTIDenQ_DEO2_FDTD_Synchronize(void *p);
const char *Prop2DAcoTTIDenQ_DEO2_FDTD_LastError(void);
void Prop2DAcoTTIDenQ_DEO2_FDTD_ClearError(void);

#ifdef __cplusplus
}
#endif

#endif
