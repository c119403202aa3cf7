// wfd_capi.hpp -- extern "C" entry points shared by the three propagator libraries.
//
// WFD_DEFINE_COMMON_CAPI(Prefix, PHYS) emits, with the reference's names and C types
// (src/prop3DAcoIsoDenQ_DEO2_FDTD.cc:6-140 and the VTI/TTI twins):
//   Prefix_alloc, _free, _SetupDtOmegaInvQ, _TimeStep, _ScaleSpatialDerivatives,
//   _getNx/_getNy/_getNz
// and the additive device-path symbols declared in include/wavefd_amd.h:
//   Prefix_SetHostCoherence / _GetHostCoherence / _SyncToDevice / _SyncToHost /
//   _Synchronize / _GetDevicePtr / _SetSources / _SetReceivers /
//   _TimeStepInjectExtract / _GetReceiverData / _SetFusedStep / _LastError
#pragma once
#include <cfloat>
#include <cstdio>
#include <cstdlib>

#include "wfd_ops.hpp"

#define WFD_CAT2(a, b) a##_##b
#define WFD_CAT(a, b) WFD_CAT2(a, b)

#define WFD_GETTER(Prefix, Name, FID)                                     \
    float *WFD_CAT(Prefix, get##Name)(void *p) {                          \
        return p ? reinterpret_cast<wfd::Prop *>(p)->host(wfd::FID) : nullptr; \
    }

#define WFD_DEFINE_COMMON_CAPI(Prefix, PHYS)                                                              \
    void *WFD_CAT(Prefix, alloc)(long fs, long nthread, long nx, long ny, long nz, long nsponge, float dx,  \
                                 float dy, float dz, float dt, long nbx, long nby, long nbz) {             \
        return (void *)wfd::make_prop(PHYS, fs, nthread, nx, ny, nz, nsponge, dx, dy, dz, dt, nbx, nby, nbz); \
    }                                                                                                     \
    void WFD_CAT(Prefix, free)(void *p) { delete reinterpret_cast<wfd::Prop *>(p); }                      \
    void WFD_CAT(Prefix, SetupDtOmegaInvQ)(void *p, float freqQ, float qMin, float qInterior) {           \
        /* src/absorb.h:98-103 reports and terminates the process; this library reports the same   */ \
        /* message, records it for LastError and returns with the handle's Q unchanged             */ \
        if (!wfd::check_freq_q(freqQ)) return;                                                            \
        wfd::setup_q(reinterpret_cast<wfd::Prop *>(p), freqQ, qMin, qInterior);                           \
    }                                                                                                     \
    void WFD_CAT(Prefix, TimeStep)(void *p) { wfd::time_step(reinterpret_cast<wfd::Prop *>(p), false); } \
    void WFD_CAT(Prefix, ScaleSpatialDerivatives)(void *p) {                                              \
        wfd::scale_spatial(reinterpret_cast<wfd::Prop *>(p));                                             \
    }                                                                                                     \
    long WFD_CAT(Prefix, getNx)(void *p) { return reinterpret_cast<wfd::Prop *>(p)->nx; }               \
    long WFD_CAT(Prefix, getNy)(void *p) { return reinterpret_cast<wfd::Prop *>(p)->ny; }               \
    long WFD_CAT(Prefix, getNz)(void *p) { return reinterpret_cast<wfd::Prop *>(p)->nz; }               \
    /* ---- additive device-path API ---- */                                                             \
    void WFD_CAT(Prefix, SetHostCoherence)(void *p, long mode) {                                          \
        wfd::Prop *pp = reinterpret_cast<wfd::Prop *>(p);                                                 \
        hipStreamSynchronize(pp->stream);                                                                 \
        pp->coherence = mode ? wfd::COHERENT : wfd::DEVICE;                                               \
    }                                                                                                     \
    long WFD_CAT(Prefix, GetHostCoherence)(void *p) { return reinterpret_cast<wfd::Prop *>(p)->coherence; } \
    void WFD_CAT(Prefix, SyncToDevice)(void *p) {                                                         \
        wfd::Prop *pp = reinterpret_cast<wfd::Prop *>(p);                                                 \
        if (pp->h2d(pp->all_mask())) hipStreamSynchronize(pp->stream);                                    \
    }                                                                                                     \
    void WFD_CAT(Prefix, SyncToHost)(void *p) {                                                           \
        wfd::Prop *pp = reinterpret_cast<wfd::Prop *>(p);                                                 \
        if (pp->d2h(pp->all_mask())) hipStreamSynchronize(pp->stream);                                    \
    }                                                                                                     \
    void WFD_CAT(Prefix, Synchronize)(void *p) {                                                          \
        hipError_t e = hipStreamSynchronize(reinterpret_cast<wfd::Prop *>(p)->stream);                    \
        if (e != hipSuccess) wfd::set_error(std::string("stream synchronize: ") + hipGetErrorString(e));  \
    }                                                                                                     \
    float *WFD_CAT(Prefix, GetDevicePtr)(void *p, const char *name) {                                     \
        wfd::Prop *pp = reinterpret_cast<wfd::Prop *>(p);                                                 \
        const int i = wfd::field_by_name(name);                                                           \
        if (i == wfd::F_Q) pp->q_profile = false, pp->q_exposed = true; /* the caller may write Q there */ \
        return (i >= 0 && pp->f[i].used) ? pp->f[i].d : nullptr;                                          \
    }                                                                                                     \
    void *WFD_CAT(Prefix, GetStream)(void *p) { return (void *)reinterpret_cast<wfd::Prop *>(p)->stream; } \
    long WFD_CAT(Prefix, SetSources)(void *p, wfd::SourcePoint32 **parts, size_t nparts, long *npts,     \
                                     float *data, size_t nt, long fields) {                               \
        return wfd::set_sources(reinterpret_cast<wfd::Prop *>(p), parts, nparts, npts, data, nt, fields) ? 0 : -1; \
    }                                                                                                     \
    long WFD_CAT(Prefix, SetReceivers)(void *p, wfd::SourcePoint32 **parts, size_t nparts, long *npts,   \
                                       size_t nt, long fields) {                                          \
        return wfd::set_receivers(reinterpret_cast<wfd::Prop *>(p), parts, nparts, npts, nt, fields) ? 0 : -1; \
    }                                                                                                     \
    long WFD_CAT(Prefix, SetSourcesInterp)(void *p, wfd::SourcePoint32 **parts, size_t nparts, long *npts, \
                                           float *data, size_t nd, float *h, size_t nh, size_t nm,          \
                                           long fields) {                                                   \
        return wfd::set_sources_interp(reinterpret_cast<wfd::Prop *>(p), parts, nparts, npts, data, nd, h, nh, \
                                       nm, fields) ? 0 : -1;                                                \
    }                                                                                                     \
    long WFD_CAT(Prefix, SetReceiversInterp)(void *p, wfd::SourcePoint32 **parts, size_t nparts, long *npts, \
                                             size_t nd, float *h, size_t nh, size_t nm, long fields) {        \
        return wfd::set_receivers_interp(reinterpret_cast<wfd::Prop *>(p), parts, nparts, npts, nd, h, nh, nm, \
                                         fields) ? 0 : -1;                                                  \
    }                                                                                                     \
    void WFD_CAT(Prefix, TimeStepInjectExtract)(void *p, long it) {                                       \
        wfd::time_step_inject_extract(reinterpret_cast<wfd::Prop *>(p), it, false);                       \
    }                                                                                                     \
    long WFD_CAT(Prefix, TimeStepInjectExtractGraph)(void *p, long it0, long nsteps) {                   \
        return wfd::time_loop_graph(reinterpret_cast<wfd::Prop *>(p), it0, nsteps) ? 0 : -1;             \
    }                                                                                                     \
    long WFD_CAT(Prefix, InjectRegistered)(void *p, long it) {                                            \
        return wfd::inject_registered_call(reinterpret_cast<wfd::Prop *>(p), it) ? 0 : -1;                \
    }                                                                                                     \
    long WFD_CAT(Prefix, ExtractRegistered)(void *p, long it) {                                           \
        return wfd::extract_registered_call(reinterpret_cast<wfd::Prop *>(p), it) ? 0 : -1;               \
    }                                                                                                     \
    long WFD_CAT(Prefix, GetReceiverData)(void *p, float *out) {                                          \
        return wfd::get_receiver_data(reinterpret_cast<wfd::Prop *>(p), out) ? 0 : -1;                    \
    }                                                                                                     \
    void WFD_CAT(Prefix, SetFusedStep)(void *p, long fused) { reinterpret_cast<wfd::Prop *>(p)->fused = (int)fused; } \
    long WFD_CAT(Prefix, SetFusedTile)(void *p, long tile) {                                              \
        return wfd::set_fused_tile(reinterpret_cast<wfd::Prop *>(p), tile) ? 0 : -1;                      \
    }                                                                                                     \
    long WFD_CAT(Prefix, GetFusedTile)(void *p) { return reinterpret_cast<wfd::Prop *>(p)->fused_tile; }  \
    long WFD_CAT(Prefix, SetFusedConfig)(void *p, long tile, long lockstep, long chunk) {                 \
        return wfd::set_fused_config(reinterpret_cast<wfd::Prop *>(p), tile, lockstep, chunk) ? 0 : -1;   \
    }                                                                                                     \
    long WFD_CAT(Prefix, GetFusedConfig)(void *p, long *out) {                                            \
        return wfd::get_fused_config(reinterpret_cast<wfd::Prop *>(p), out) ? 0 : -1;                     \
    }                                                                                                     \
    long WFD_CAT(Prefix, TimeStepBatch)(void **ps, long n, long epilogue, float **a0, float **a1, float **a2) { \
        return wfd::time_step_batch(PHYS, ps, n, epilogue, a0, a1, a2) ? 0 : -1;                          \
    }                                                                                                     \
    const char *WFD_CAT(Prefix, LastError)(void) { return wfd::g_last_error.c_str(); }                    \
    void WFD_CAT(Prefix, ClearError)(void) { wfd::g_last_error.clear(); }
