// libprop3DAcoVTIDenQ_DEO2_FDTD -- MI355X (gfx950) drop-in for the reference library of the
// same name (src/prop3DAcoVTIDenQ_DEO2_FDTD.cc, built by src/CMakeLists.txt:11).  Declarations
// and citations are in include/prop3DAcoVTIDenQ_DEO2_FDTD.h.
#include "wfd_capi.hpp"

#define VTI_PREFIX Prop3DAcoVTIDenQ_DEO2_FDTD

#pragma GCC visibility push(default)
extern "C" {

WFD_DEFINE_COMMON_CAPI(VTI_PREFIX, wfd::VTI)

void Prop3DAcoVTIDenQ_DEO2_FDTD_TimeStepLinear(void *p) { wfd::time_step(reinterpret_cast<wfd::Prop *>(p), true); }

void Prop3DAcoVTIDenQ_DEO2_FDTD_ForwardBornInjection_V(void *p, float *dVel, float *wavefieldDP, float *wavefieldDM) {
    wfd::vti_inj_v(reinterpret_cast<wfd::Prop *>(p), dVel, wavefieldDP, wavefieldDM);
}

void Prop3DAcoVTIDenQ_DEO2_FDTD_ForwardBornInjection_VEA(void *p, float *dVel, float *dEps, float *dEta,
                                                         float *wavefieldP, float *wavefieldM, float *wavefieldDP,
                                                         float *wavefieldDM) {
    wfd::vea_inj(reinterpret_cast<wfd::Prop *>(p), dVel, dEps, dEta, wavefieldP, wavefieldM, wavefieldDP,
                 wavefieldDM);
}

void Prop3DAcoVTIDenQ_DEO2_FDTD_AdjointBornAccumulation_V(void *p, float *dVel, float *wavefieldDP,
                                                          float *wavefieldDM) {
    wfd::vti_acc_v(reinterpret_cast<wfd::Prop *>(p), dVel, wavefieldDP, wavefieldDM);
}

void Prop3DAcoVTIDenQ_DEO2_FDTD_AdjointBornAccumulation_wavefieldsep_V(void *p, float *dVel, float *wavefieldDP,
                                                                       float *wavefieldDM, const long isFWI) {
    wfd::wavefield_sep(reinterpret_cast<wfd::Prop *>(p), dVel, wavefieldDP, wavefieldDM, isFWI);
}

void Prop3DAcoVTIDenQ_DEO2_FDTD_AdjointBornAccumulation_VEA(void *p, float *dVel, float *dEps, float *dEta,
                                                            float *wavefieldP, float *wavefieldM, float *wavefieldDP,
                                                            float *wavefieldDM) {
    wfd::vea_acc(reinterpret_cast<wfd::Prop *>(p), dVel, dEps, dEta, wavefieldP, wavefieldM, wavefieldDP,
                 wavefieldDM);
}

// additive: the Born loops' per-step pairs as one fused step (include/wavefd_amd.h)
long Prop3DAcoVTIDenQ_DEO2_FDTD_TimeStepSaveDP(void *p, float *DP, float *DM) {
    return wfd::born_step(reinterpret_cast<wfd::Prop *>(p), wfd::EPI_SAVE_DP, DP, DM, nullptr) ? 0 : -1;
}

long Prop3DAcoVTIDenQ_DEO2_FDTD_TimeStepBornInjection_V(void *p, float *dVel, float *DP, float *DM) {
    return wfd::born_step(reinterpret_cast<wfd::Prop *>(p), wfd::EPI_BORN_INJ, dVel, DP, DM) ? 0 : -1;
}

long Prop3DAcoVTIDenQ_DEO2_FDTD_TimeStepAdjointBornAccumulation_V(void *p, float *dVel, float *DP, float *DM) {
    return wfd::born_step(reinterpret_cast<wfd::Prop *>(p), wfd::EPI_IMAGE, dVel, DP, DM) ? 0 : -1;
}

WFD_GETTER(VTI_PREFIX, V, F_V)
WFD_GETTER(VTI_PREFIX, Eps, F_EPS)
WFD_GETTER(VTI_PREFIX, Eta, F_ETA)
WFD_GETTER(VTI_PREFIX, B, F_B)
WFD_GETTER(VTI_PREFIX, F, F_F)
WFD_GETTER(VTI_PREFIX, PSpace, F_PSPACE)
WFD_GETTER(VTI_PREFIX, MSpace, F_MSPACE)
WFD_GETTER(VTI_PREFIX, PCur, F_PCUR)
WFD_GETTER(VTI_PREFIX, POld, F_POLD)
WFD_GETTER(VTI_PREFIX, MCur, F_MCUR)
WFD_GETTER(VTI_PREFIX, MOld, F_MOLD)
// Not exported by the reference's shim although its Julia wrapper generates DtOmegaInvQ(prop)
// for it (src/prop3DAcoVTIDenQ_DEO2_FDTD.jl:58); exported here so that wrapper works.
WFD_GETTER(VTI_PREFIX, DtOmegaInvQ, F_Q)

}  // extern "C"
#pragma GCC visibility pop
