// libprop2DAcoIsoDenQ_DEO2_FDTD -- MI355X (gfx950) drop-in for the reference library of the
// same name (src/prop2DAcoIsoDenQ_DEO2_FDTD.cc).  Entry points keep the reference's names,
// argument order and C types; declarations and citations are in
// include/prop2DAcoIsoDenQ_DEO2_FDTD.h.  The handle is the 3D one with ny = 1 (wfd_2d.hpp).
#include "wfd_ops.hpp"

#pragma GCC visibility push(default)
extern "C" {

#define P2(p) reinterpret_cast<wfd::Prop *>(p)

void *Prop2DAcoIsoDenQ_DEO2_FDTD_alloc(long fs, long nthread, long nx, long nz, long nsponge, float dx, float dz,
                                       float dt, long nbx, long nbz) {
    return (void *)wfd::make_prop_2d(wfd::ISO, fs, nthread, nx, nz, nsponge, dx, dz, dt, nbx, nbz);
}

void Prop2DAcoIsoDenQ_DEO2_FDTD_free(void *p) { delete P2(p); }

void Prop2DAcoIsoDenQ_DEO2_FDTD_SetupDtOmegaInvQ(void *p, float freqQ, float qMin, float qInterior) {
    // src/absorb.h:34-39 exits the process; this reports, records LastError and returns
    if (!wfd::check_freq_q(freqQ)) return;
    wfd::setup_q_2d(P2(p), freqQ, qMin, qInterior);
}

void Prop2DAcoIsoDenQ_DEO2_FDTD_TimeStep(void *p) { wfd::time_step(P2(p), false); }

void Prop2DAcoIsoDenQ_DEO2_FDTD_ScaleSpatialDerivatives(void *p) { wfd::scale_spatial(P2(p)); }

void Prop2DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_V(void *p, float *dVel, float *wavefieldDP) {
    wfd::iso2_inj_v(P2(p), dVel, wavefieldDP);
}

void Prop2DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_VB(void *p, float *dVel, float *dBuoy, float *wavefieldP,
                                                        float *wavefieldDP) {
    wfd::iso2_inj_vb(P2(p), dVel, dBuoy, wavefieldP, wavefieldDP);
}

void Prop2DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_B(void *p, float *dBuoy, float *wavefieldP, float *wavefieldDP) {
    wfd::iso2_inj_vb(P2(p), nullptr, dBuoy, wavefieldP, wavefieldDP);
}

void Prop2DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_V(void *p, float *dVel, float *wavefieldDP) {
    wfd::iso2_acc(P2(p), dVel, nullptr, nullptr, wavefieldDP);
}

void Prop2DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_VB(void *p, float *dVel, float *dBuoy, float *wavefieldP,
                                                           float *wavefieldDP) {
    wfd::iso2_acc(P2(p), dVel, dBuoy, wavefieldP, wavefieldDP);
}

void Prop2DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_B(void *p, float *dBuoy, float *wavefieldP,
                                                          float *wavefieldDP) {
    wfd::iso2_acc(P2(p), nullptr, dBuoy, wavefieldP, wavefieldDP);
}

void Prop2DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_wavefieldsep(void *p, float *dVel, float *wavefieldDP,
                                                                     const long isFWI) {
    wfd::wavefield_sep(P2(p), dVel, wavefieldDP, nullptr, isFWI);
}

long Prop2DAcoIsoDenQ_DEO2_FDTD_getNx(void *p) { return P2(p)->nx; }
long Prop2DAcoIsoDenQ_DEO2_FDTD_getNz(void *p) { return P2(p)->nz; }
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getV(void *p) { return P2(p)->host(wfd::F_V); }
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getB(void *p) { return P2(p)->host(wfd::F_B); }
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getDtOmegaInvQ(void *p) { return P2(p)->host(wfd::F_Q); }
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getPSpace(void *p) { return P2(p)->host(wfd::F_PSPACE); }
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getPCur(void *p) { return P2(p)->host(wfd::F_PCUR); }
float *Prop2DAcoIsoDenQ_DEO2_FDTD_getPOld(void *p) { return P2(p)->host(wfd::F_POLD); }

// additive (include/prop2DAcoIsoDenQ_DEO2_FDTD.h): host coherence, device addresses, errors
void Prop2DAcoIsoDenQ_DEO2_FDTD_SetHostCoherence(void *p, long mode) {
    (void)hipStreamSynchronize(P2(p)->stream);
    P2(p)->coherence = mode ? wfd::COHERENT : wfd::DEVICE;
}
void Prop2DAcoIsoDenQ_DEO2_FDTD_SyncToDevice(void *p) {
    if (P2(p)->h2d(P2(p)->all_mask())) (void)hipStreamSynchronize(P2(p)->stream);
}
void Prop2DAcoIsoDenQ_DEO2_FDTD_SyncToHost(void *p) {
    if (P2(p)->d2h(P2(p)->all_mask())) (void)hipStreamSynchronize(P2(p)->stream);
}
float *Prop2DAcoIsoDenQ_DEO2_FDTD_GetDevicePtr(void *p, const char *name) {
    const int i = wfd::field_by_name(name);
    if (i < 0 || !P2(p)->f[i].used) return nullptr;
    if (i == wfd::F_Q) P2(p)->q_profile = false, P2(p)->q_exposed = true;  // the caller may write Q there
    return P2(p)->f[i].d;
}
void Prop2DAcoIsoDenQ_DEO2_FDTD_SetFusedStep(void *p, long fused) { P2(p)->fused = fused ? 1 : 0; }
// additive: the Born loops' per-step pairs as one step (same results as the two calls)
long Prop2DAcoIsoDenQ_DEO2_FDTD_TimeStepSaveDP(void *p, float *DP) {
    return wfd::born_step(P2(p), wfd::EPI_SAVE_DP, DP, nullptr, nullptr) ? 0 : -1;
}
long Prop2DAcoIsoDenQ_DEO2_FDTD_TimeStepBornInjection_V(void *p, float *dVel, float *DP) {
    return wfd::born_step(P2(p), wfd::EPI_BORN_INJ, dVel, DP, nullptr) ? 0 : -1;
}
long Prop2DAcoIsoDenQ_DEO2_FDTD_TimeStepAdjointBornAccumulation_V(void *p, float *dVel, float *DP) {
    return wfd::born_step(P2(p), wfd::EPI_IMAGE, dVel, DP, nullptr) ? 0 : -1;
}
void *Prop2DAcoIsoDenQ_DEO2_FDTD_GetStream(void *p) { return (void *)P2(p)->stream; }
void Prop2DAcoIsoDenQ_DEO2_FDTD_Synchronize(void *p) {
    if (hipStreamSynchronize(P2(p)->stream) != hipSuccess) wfd::set_error("Synchronize: stream error");
}
const char *Prop2DAcoIsoDenQ_DEO2_FDTD_LastError(void) { return wfd::g_last_error.c_str(); }
void Prop2DAcoIsoDenQ_DEO2_FDTD_ClearError(void) { wfd::g_last_error.clear(); }

}  // extern "C"
#pragma GCC visibility pop
