// libprop2DAcoVTIDenQ_DEO2_FDTD -- MI355X (gfx950) drop-in for the reference library of the
// same name (src/prop2DAcoVTIDenQ_DEO2_FDTD.cc).  Entry points keep the reference's names,
// argument order and C types; declarations and citations are in
// include/prop2DAcoVTIDenQ_DEO2_FDTD.h.  The handle is the 3D one with ny = 1 (wfd_2d.hpp).
#include "wfd_ops.hpp"

#pragma GCC visibility push(default)
extern "C" {

#define P2(p) reinterpret_cast<wfd::Prop *>(p)

void *Prop2DAcoVTIDenQ_DEO2_FDTD_alloc(long fs, long nthread, long nx, long nz, long nsponge, float dx, float dz, float dt, long nbx,
              long nbz) {
    return (void *)wfd::make_prop_2d(wfd::VTI, fs, nthread, nx, nz, nsponge, dx, dz, dt, nbx, nbz);
}

void Prop2DAcoVTIDenQ_DEO2_FDTD_free(void *p) { delete P2(p); }

void Prop2DAcoVTIDenQ_DEO2_FDTD_SetupDtOmegaInvQ(void *p, float freqQ, float qMin, float qInterior) {
    // src/absorb.h:34-39 exits the process; this reports, records LastError and returns
    if (!wfd::check_freq_q(freqQ)) return;
    wfd::setup_q_2d(P2(p), freqQ, qMin, qInterior);
}

void Prop2DAcoVTIDenQ_DEO2_FDTD_TimeStep(void *p) { wfd::time_step(P2(p), false); }

void Prop2DAcoVTIDenQ_DEO2_FDTD_TimeStepLinear(void *p) { wfd::time_step(P2(p), true); }

void Prop2DAcoVTIDenQ_DEO2_FDTD_ScaleSpatialDerivatives(void *p) { wfd::scale_spatial(P2(p)); }

void Prop2DAcoVTIDenQ_DEO2_FDTD_ForwardBornInjection_V(void *p, float *dVel, float *wavefieldDP, float *wavefieldDM) {
    wfd::vti_inj_v(P2(p), dVel, wavefieldDP, wavefieldDM);
}

void Prop2DAcoVTIDenQ_DEO2_FDTD_ForwardBornInjection_VEA(void *p, float *dVel, float *dEps, float *dEta, float *wavefieldP,
                                  float *wavefieldM, float *wavefieldDP, float *wavefieldDM) {
    wfd::aniso2_inj_vea(P2(p), dVel, dEps, dEta, wavefieldP, wavefieldM, wavefieldDP, wavefieldDM);
}

void Prop2DAcoVTIDenQ_DEO2_FDTD_AdjointBornAccumulation_V(void *p, float *dVel, float *wavefieldDP, float *wavefieldDM) {
    wfd::vti_acc_v(P2(p), dVel, wavefieldDP, wavefieldDM);
}

void Prop2DAcoVTIDenQ_DEO2_FDTD_AdjointBornAccumulation_wavefieldsep_V(void *p, float *dVel, float *wavefieldDP, float *wavefieldDM,
                                                const long isFWI) {
    wfd::wavefield_sep(P2(p), dVel, wavefieldDP, wavefieldDM, isFWI);
}

void Prop2DAcoVTIDenQ_DEO2_FDTD_AdjointBornAccumulation_VEA(void *p, float *dVel, float *dEps, float *dEta, float *wavefieldP,
                                     float *wavefieldM, float *wavefieldDP, float *wavefieldDM) {
    wfd::aniso2_acc_vea(P2(p), dVel, dEps, dEta, wavefieldP, wavefieldM, wavefieldDP, wavefieldDM);
}

long Prop2DAcoVTIDenQ_DEO2_FDTD_getNx(void *p) { return P2(p)->nx; }
long Prop2DAcoVTIDenQ_DEO2_FDTD_getNz(void *p) { return P2(p)->nz; }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getV(void *p) { return P2(p)->host(wfd::F_V); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getEps(void *p) { return P2(p)->host(wfd::F_EPS); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getEta(void *p) { return P2(p)->host(wfd::F_ETA); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getB(void *p) { return P2(p)->host(wfd::F_B); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getF(void *p) { return P2(p)->host(wfd::F_F); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getDtOmegaInvQ(void *p) { return P2(p)->host(wfd::F_Q); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getPSpace(void *p) { return P2(p)->host(wfd::F_PSPACE); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getMSpace(void *p) { return P2(p)->host(wfd::F_MSPACE); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getPCur(void *p) { return P2(p)->host(wfd::F_PCUR); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getPOld(void *p) { return P2(p)->host(wfd::F_POLD); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getMCur(void *p) { return P2(p)->host(wfd::F_MCUR); }
float *Prop2DAcoVTIDenQ_DEO2_FDTD_getMOld(void *p) { return P2(p)->host(wfd::F_MOLD); }

void Prop2DAcoVTIDenQ_DEO2_FDTD_applyFirstDerivatives2D_PlusHalf(const long freeSurface, const long nx, const long nz, const long nthread,
                                          const float c8_1, const float c8_2, const float c8_3, const float c8_4,
                                          const float invDx, const float invDz, const float *__restrict__ const inX,
                                          const float *__restrict__ const inZ, float *__restrict__ outX,
                                          float *__restrict__ outZ, const long BX_2D, const long BZ_2D) {
    (void)nthread;
    (void)BX_2D;
    (void)BZ_2D;
    wfd::dplus_2d_host(freeSurface, nx, nz, c8_1, c8_2, c8_3, c8_4, invDx, invDz, inX, inZ, outX, outZ);
}

// additive (include/prop2DAcoVTIDenQ_DEO2_FDTD.h): host coherence, device addresses, errors
void Prop2DAcoVTIDenQ_DEO2_FDTD_SetHostCoherence(void *p, long mode) {
    (void)hipStreamSynchronize(P2(p)->stream);
    P2(p)->coherence = mode ? wfd::COHERENT : wfd::DEVICE;
}
void Prop2DAcoVTIDenQ_DEO2_FDTD_SyncToDevice(void *p) {
    if (P2(p)->h2d(P2(p)->all_mask())) (void)hipStreamSynchronize(P2(p)->stream);
}
void Prop2DAcoVTIDenQ_DEO2_FDTD_SyncToHost(void *p) {
    if (P2(p)->d2h(P2(p)->all_mask())) (void)hipStreamSynchronize(P2(p)->stream);
}
float *Prop2DAcoVTIDenQ_DEO2_FDTD_GetDevicePtr(void *p, const char *name) {
    const int i = wfd::field_by_name(name);
    if (i < 0 || !P2(p)->f[i].used) return nullptr;
    if (i == wfd::F_Q) P2(p)->q_profile = false, P2(p)->q_exposed = true;  // the caller may write Q there
    return P2(p)->f[i].d;
}
void Prop2DAcoVTIDenQ_DEO2_FDTD_SetFusedStep(void *p, long fused) { P2(p)->fused = fused ? 1 : 0; }
// additive: the Born loops' per-step pairs as one step (same results as the two calls)
long Prop2DAcoVTIDenQ_DEO2_FDTD_TimeStepSaveDP(void *p, float *DP, float *DM) {
    return wfd::born_step(P2(p), wfd::EPI_SAVE_DP, DP, DM, nullptr) ? 0 : -1;
}
long Prop2DAcoVTIDenQ_DEO2_FDTD_TimeStepBornInjection_V(void *p, float *dVel, float *DP, float *DM) {
    return wfd::born_step(P2(p), wfd::EPI_BORN_INJ, dVel, DP, DM) ? 0 : -1;
}
long Prop2DAcoVTIDenQ_DEO2_FDTD_TimeStepAdjointBornAccumulation_V(void *p, float *dVel, float *DP, float *DM) {
    return wfd::born_step(P2(p), wfd::EPI_IMAGE, dVel, DP, DM) ? 0 : -1;
}
void *Prop2DAcoVTIDenQ_DEO2_FDTD_GetStream(void *p) { return (void *)P2(p)->stream; }
void Prop2DAcoVTIDenQ_DEO2_FDTD_Synchronize(void *p) {
    if (hipStreamSynchronize(P2(p)->stream) != hipSuccess) wfd::set_error("Synchronize: stream error");
}
const char *Prop2DAcoVTIDenQ_DEO2_FDTD_LastError(void) { return wfd::g_last_error.c_str(); }
void Prop2DAcoVTIDenQ_DEO2_FDTD_ClearError(void) { wfd::g_last_error.clear(); }

}  // extern "C"
#pragma GCC visibility pop
