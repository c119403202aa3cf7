/*
 * libprop3DAcoTTIDenQ_DEO2_FDTD -- C-ABI of the MI355X (gfx950) TTI variable-density
 * visco-acoustic propagator.  Drop-in for the reference library of the same name
 * (src/prop3DAcoTTIDenQ_DEO2_FDTD.cc, target in src/CMakeLists.txt:11).  Layout and ownership
 * as in prop3DAcoIsoDenQ_DEO2_FDTD.h; additive symbols in wavefd_amd.h.
 */
#ifndef WAVEFD_AMD_PROP3DACOTTIDENQ_DEO2_FDTD_H
#define WAVEFD_AMD_PROP3DACOTTIDENQ_DEO2_FDTD_H

#ifdef __cplusplus
extern "C" {
#endif

/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:6-26 */
void *Prop3DAcoTTIDenQ_DEO2_FDTD_alloc(long fs, long nthread, long nx, long ny, long nz, long nsponge, float dx,
                                       float dy, float dz, float dt, long nbx, long nby, long nbz);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:28-31 */
void Prop3DAcoTTIDenQ_DEO2_FDTD_free(void *p);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:33-43 */
void Prop3DAcoTTIDenQ_DEO2_FDTD_SetupDtOmegaInvQ(void *p, float freqQ, float qMin, float qInterior);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:45-48 (h:359-379) */
void Prop3DAcoTTIDenQ_DEO2_FDTD_TimeStep(void *p);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:50-53 (h:387-410) */
void Prop3DAcoTTIDenQ_DEO2_FDTD_ScaleSpatialDerivatives(void *p);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:55-59 (h:423-455) */
void Prop3DAcoTTIDenQ_DEO2_FDTD_ForwardBornInjection_V(void *p, float *dVel, float *wavefieldDP, float *wavefieldDM);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:61-66 (h:460-554) */
void Prop3DAcoTTIDenQ_DEO2_FDTD_ForwardBornInjection_VEA(void *p, float *dVel, float *dEps, float *dEta,
                                                         float *wavefieldP, float *wavefieldM, float *wavefieldDP,
                                                         float *wavefieldDM);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:68-72 (h:567-592) */
void Prop3DAcoTTIDenQ_DEO2_FDTD_AdjointBornAccumulation_V(void *p, float *dVel, float *wavefieldDP,
                                                          float *wavefieldDM);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:74-78 (h:609-730) */
void Prop3DAcoTTIDenQ_DEO2_FDTD_AdjointBornAccumulation_wavefieldsep_V(void *p, float *dVel, float *wavefieldDP,
                                                                       float *wavefieldDM, const long isFWI);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:80-85 (h:733-790) */
void Prop3DAcoTTIDenQ_DEO2_FDTD_AdjointBornAccumulation_VEA(void *p, float *dVel, float *dEps, float *dEta,
                                                            float *wavefieldP, float *wavefieldM, float *wavefieldDP,
                                                            float *wavefieldDM);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:87-100 */
long Prop3DAcoTTIDenQ_DEO2_FDTD_getNx(void *p);
long Prop3DAcoTTIDenQ_DEO2_FDTD_getNy(void *p);
long Prop3DAcoTTIDenQ_DEO2_FDTD_getNz(void *p);
/* src/prop3DAcoTTIDenQ_DEO2_FDTD.cc:102-181 */
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getV(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getEps(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getEta(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getSinTheta(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getCosTheta(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getSinPhi(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getCosPhi(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getB(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getF(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getDtOmegaInvQ(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getPSpace(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getMSpace(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getPCur(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getPOld(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getMCur(void *p);
float *Prop3DAcoTTIDenQ_DEO2_FDTD_getMOld(void *p);

#ifdef __cplusplus
}
#endif
#endif
