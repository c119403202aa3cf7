/*
 * libprop3DAcoVTIDenQ_DEO2_FDTD -- C-ABI of the MI355X (gfx950) VTI variable-density
 * visco-acoustic propagator (coupled P/M fields).  Drop-in for the reference library of the
 * same name (src/prop3DAcoVTIDenQ_DEO2_FDTD.cc, target in src/CMakeLists.txt:11).  Layout and
 * ownership as in prop3DAcoIsoDenQ_DEO2_FDTD.h; additive symbols in wavefd_amd.h.
 */
#ifndef WAVEFD_AMD_PROP3DACOVTIDENQ_DEO2_FDTD_H
#define WAVEFD_AMD_PROP3DACOVTIDENQ_DEO2_FDTD_H

#ifdef __cplusplus
extern "C" {
#endif

/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:6-26 */
void *Prop3DAcoVTIDenQ_DEO2_FDTD_alloc(long fs, long nthread, long nx, long ny, long nz, long nsponge, float dx,
                                       float dy, float dz, float dt, long nbx, long nby, long nbz);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:28-31 */
void Prop3DAcoVTIDenQ_DEO2_FDTD_free(void *p);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:33-43 */
void Prop3DAcoVTIDenQ_DEO2_FDTD_SetupDtOmegaInvQ(void *p, float freqQ, float qMin, float qInterior);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:45-48 (h:356-376): step without pSpace/mSpace output */
void Prop3DAcoVTIDenQ_DEO2_FDTD_TimeStepLinear(void *p);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:50-53 (h:326-346) */
void Prop3DAcoVTIDenQ_DEO2_FDTD_TimeStep(void *p);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:55-58 (h:384-407) */
void Prop3DAcoVTIDenQ_DEO2_FDTD_ScaleSpatialDerivatives(void *p);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:60-64 (h:420-452) */
void Prop3DAcoVTIDenQ_DEO2_FDTD_ForwardBornInjection_V(void *p, float *dVel, float *wavefieldDP, float *wavefieldDM);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:66-71 (h:457-551) */
void Prop3DAcoVTIDenQ_DEO2_FDTD_ForwardBornInjection_VEA(void *p, float *dVel, float *dEps, float *dEta,
                                                         float *wavefieldP, float *wavefieldM, float *wavefieldDP,
                                                         float *wavefieldDM);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:73-77 (h:564-589) */
void Prop3DAcoVTIDenQ_DEO2_FDTD_AdjointBornAccumulation_V(void *p, float *dVel, float *wavefieldDP,
                                                          float *wavefieldDM);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:79-83 (h:606-725) */
void Prop3DAcoVTIDenQ_DEO2_FDTD_AdjointBornAccumulation_wavefieldsep_V(void *p, float *dVel, float *wavefieldDP,
                                                                       float *wavefieldDM, const long isFWI);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:85-90 (h:730-787) */
void Prop3DAcoVTIDenQ_DEO2_FDTD_AdjointBornAccumulation_VEA(void *p, float *dVel, float *dEps, float *dEta,
                                                            float *wavefieldP, float *wavefieldM, float *wavefieldDP,
                                                            float *wavefieldDM);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:92-105 */
long Prop3DAcoVTIDenQ_DEO2_FDTD_getNx(void *p);
long Prop3DAcoVTIDenQ_DEO2_FDTD_getNy(void *p);
long Prop3DAcoVTIDenQ_DEO2_FDTD_getNz(void *p);
/* src/prop3DAcoVTIDenQ_DEO2_FDTD.cc:107-161 */
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getV(void *p);
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getEps(void *p);
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getEta(void *p);
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getB(void *p);
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getF(void *p);
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getPSpace(void *p);
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getMSpace(void *p);
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getPCur(void *p);
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getPOld(void *p);
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getMCur(void *p);
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getMOld(void *p);
/* Not exported by the reference's shim, but its Julia wrapper generates DtOmegaInvQ(prop) for
 * it (src/prop3DAcoVTIDenQ_DEO2_FDTD.jl:58); exported here so that accessor works. */
float *Prop3DAcoVTIDenQ_DEO2_FDTD_getDtOmegaInvQ(void *p);

#ifdef __cplusplus
}
#endif
#endif
