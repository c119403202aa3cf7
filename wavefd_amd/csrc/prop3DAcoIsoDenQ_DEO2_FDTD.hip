// libprop3DAcoIsoDenQ_DEO2_FDTD -- MI355X (gfx950) drop-in for the reference library of the
// same name (src/prop3DAcoIsoDenQ_DEO2_FDTD.cc, built by src/CMakeLists.txt:11).  Entry points
// keep the reference's names, argument order and C types; declarations and citations are in
// include/prop3DAcoIsoDenQ_DEO2_FDTD.h.
#include "wfd_capi.hpp"

#define ISO_PREFIX Prop3DAcoIsoDenQ_DEO2_FDTD

#pragma GCC visibility push(default)
extern "C" {

WFD_DEFINE_COMMON_CAPI(ISO_PREFIX, wfd::ISO)

void Prop3DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_V(void *p, float *dVel, float *wavefieldDP) {
    wfd::iso_inj_v(reinterpret_cast<wfd::Prop *>(p), dVel, wavefieldDP);
}

void Prop3DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_VB(void *p, float *dVel, float *dBuoy, float *wavefieldP,
                                                        float *wavefieldDP) {
    wfd::iso_inj_vb(reinterpret_cast<wfd::Prop *>(p), dVel, dBuoy, wavefieldP, wavefieldDP);
}

void Prop3DAcoIsoDenQ_DEO2_FDTD_ForwardBornInjection_B(void *p, float *dBuoy, float *wavefieldP, float *wavefieldDP) {
    wfd::iso_inj_vb(reinterpret_cast<wfd::Prop *>(p), nullptr, dBuoy, wavefieldP, wavefieldDP);
}

void Prop3DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_V(void *p, float *dVel, float *wavefieldDP) {
    wfd::iso_acc_v(reinterpret_cast<wfd::Prop *>(p), dVel, wavefieldDP);
}

void Prop3DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_VB(void *p, float *dVel, float *dBuoy, float *wavefieldP,
                                                           float *wavefieldDP) {
    wfd::iso_acc_vb(reinterpret_cast<wfd::Prop *>(p), dVel, dBuoy, wavefieldP, wavefieldDP);
}

void Prop3DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_B(void *p, float *dBuoy, float *wavefieldP,
                                                          float *wavefieldDP) {
    wfd::iso_acc_vb(reinterpret_cast<wfd::Prop *>(p), nullptr, dBuoy, wavefieldP, wavefieldDP);
}

void Prop3DAcoIsoDenQ_DEO2_FDTD_AdjointBornAccumulation_wavefieldsep(void *p, float *dVel, float *wavefieldDP,
                                                                     const long isFWI) {
    wfd::wavefield_sep(reinterpret_cast<wfd::Prop *>(p), dVel, wavefieldDP, nullptr, isFWI);
}

// additive: the Born loops' per-step pairs as one fused step (include/wavefd_amd.h)
long Prop3DAcoIsoDenQ_DEO2_FDTD_TimeStepSaveDP(void *p, float *DP) {
    return wfd::born_step(reinterpret_cast<wfd::Prop *>(p), wfd::EPI_SAVE_DP, DP, nullptr, nullptr) ? 0 : -1;
}

long Prop3DAcoIsoDenQ_DEO2_FDTD_TimeStepBornInjection_V(void *p, float *dVel, float *DP) {
    return wfd::born_step(reinterpret_cast<wfd::Prop *>(p), wfd::EPI_BORN_INJ, dVel, DP, nullptr) ? 0 : -1;
}

long Prop3DAcoIsoDenQ_DEO2_FDTD_TimeStepAdjointBornAccumulation_V(void *p, float *dVel, float *DP) {
    return wfd::born_step(reinterpret_cast<wfd::Prop *>(p), wfd::EPI_IMAGE, dVel, DP, nullptr) ? 0 : -1;
}

long Prop3DAcoIsoDenQ_DEO2_FDTD_TimeStepBornCoupled_V(void *p, float *dVel, float *DP) {
    return wfd::born_coupled(reinterpret_cast<wfd::Prop *>(p), dVel, DP) ? 0 : -1;
}

WFD_GETTER(ISO_PREFIX, V, F_V)
WFD_GETTER(ISO_PREFIX, B, F_B)
WFD_GETTER(ISO_PREFIX, DtOmegaInvQ, F_Q)
WFD_GETTER(ISO_PREFIX, PSpace, F_PSPACE)
WFD_GETTER(ISO_PREFIX, PCur, F_PCUR)
WFD_GETTER(ISO_PREFIX, POld, F_POLD)

}  // extern "C"
#pragma GCC visibility pop
