/*
 * libprop3DAcoIsoDenQ_DEO2_FDTD -- C-ABI of the MI355X (gfx950) isotropic variable-density
 * visco-acoustic propagator.  Drop-in for the reference library of the same name
 * (src/prop3DAcoIsoDenQ_DEO2_FDTD.cc, target in src/CMakeLists.txt:11).  Every reference
 * symbol keeps its name, argument order and C types; each is listed with the reference line
 * it replaces.  Additive symbols (no reference counterpart) are in wavefd_amd.h.
 *
 * Arrays: nx*ny*nz float32, index kx*ny*nz + ky*nz + kz (z fastest) -- Julia's (nz,ny,nx).
 * Getters return host pointers owned by the library (valid until _free); see wavefd_amd.h for
 * how host arrays and device state are kept coherent.
 */
#ifndef WAVEFD_AMD_PROP3DACOISODENQ_DEO2_FDTD_H
#define WAVEFD_AMD_PROP3DACOISODENQ_DEO2_FDTD_H

#ifdef __cplusplus
extern "C" {
#endif

/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:6-26 -- returns NULL (and sets the last error) on failure */
void *Prop3DAcoIsoDenQ_DEO2_FDTD_alloc(long fs, long nthread, long nx, long ny, long nz, long nsponge, float dx,
                                       float dy, float dz, float dt, long nbx, long nby, long nbz);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:28-31 */
void Prop3DAcoIsoDenQ_DEO2_FDTD_free(void *p);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:33-43 (src/absorb.h:85-151); exits the process when
 * freqQ < FLT_EPSILON, as the reference does (src/absorb.h:98-103) */
void Prop3DAcoIsoDenQ_DEO2_FDTD_SetupDtOmegaInvQ(void *p, float freqQ, float qMin, float qInterior);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:45-48 (h:247-261): one step, then pCur/pOld swap */
void Prop3DAcoIsoDenQ_DEO2_FDTD_TimeStep(void *p);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:50-53 (h:269-291) */
void Prop3DAcoIsoDenQ_DEO2_FDTD_ScaleSpatialDerivatives(void *p);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:55-59 (h:305-334) */
void Prop3DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_V(void *p, float *dVel, float *wavefieldDP);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:61-65 (h:351-410) */
void Prop3DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_VB(void *p, float *dVel, float *dBuoy, float *wavefieldP,
                                                        float *wavefieldDP);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:67-71 (h:424-498) */
void Prop3DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_B(void *p, float *dBuoy, float *wavefieldP, float *wavefieldDP);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:73-77 (h:518-546) */
void Prop3DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_V(void *p, float *dVel, float *wavefieldDP);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:79-83 (h:563-605) */
void Prop3DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_VB(void *p, float *dVel, float *dBuoy, float *wavefieldP,
                                                           float *wavefieldDP);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:85-89 (h:618-656) */
void Prop3DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_B(void *p, float *dBuoy, float *wavefieldP,
                                                          float *wavefieldDP);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:91-95 (h:673-765) */
void Prop3DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_wavefieldsep(void *p, float *dVel, float *wavefieldDP,
                                                                     const long isFWI);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:97-110 */
long Prop3DAcoIsoDenQ_DEO2_FDTD_getNx(void *p);
long Prop3DAcoIsoDenQ_DEO2_FDTD_getNy(void *p);
long Prop3DAcoIsoDenQ_DEO2_FDTD_getNz(void *p);
/* src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:112-140 */
float *Prop3DAcoIsoDenQ_DEO2_FDTD_getV(void *p);
float *Prop3DAcoIsoDenQ_DEO2_FDTD_getB(void *p);
float *Prop3DAcoIsoDenQ_DEO2_FDTD_getDtOmegaInvQ(void *p);
float *Prop3DAcoIsoDenQ_DEO2_FDTD_getPSpace(void *p);
float *Prop3DAcoIsoDenQ_DEO2_FDTD_getPCur(void *p);
float *Prop3DAcoIsoDenQ_DEO2_FDTD_getPOld(void *p);

#ifdef __cplusplus
}
#endif
#endif
