"""wavefd_amd -- MI355X (gfx950) implementation of WaveFD's acoustic propagator path.

The compute lives in in-tree HIP libraries with the reference's names and C-ABI
(wavefd_amd/lib/lib{prop3DAco{Iso,VTI,TTI}DenQ_DEO2_FDTD,prop2DAco{Iso,VTI,TTI}DenQ_DEO2_FDTD,
spacetime}.so); this package is the host-side mirror of the reference's Julia API over them
(the 2D propagators in wavefd_amd.prop2d).
"""
from .prop3d import (  # noqa: F401
    B, CosPhi, CosTheta, DtOmegaInvQ, Eps, Eta, F, ImagingConditionStandard,
    ImagingConditionWaveFieldSeparationFWI, ImagingConditionWaveFieldSeparationMIX,
    ImagingConditionWaveFieldSeparationRTM, MCur, MOld, MSpace, PCur, POld, PSpace,
    Prop3DAcoIsoDenQ_DEO2_FDTD, Prop3DAcoIsoDenQ_DEO2_FDTD_Model_B, Prop3DAcoIsoDenQ_DEO2_FDTD_Model_V,
    Prop3DAcoIsoDenQ_DEO2_FDTD_Model_VB, Prop3DAcoTTIDenQ_DEO2_FDTD, Prop3DAcoTTIDenQ_DEO2_FDTD_Model_V,
    Prop3DAcoTTIDenQ_DEO2_FDTD_Model_VEA, Prop3DAcoVTIDenQ_DEO2_FDTD, Prop3DAcoVTIDenQ_DEO2_FDTD_Model_V,
    Prop3DAcoVTIDenQ_DEO2_FDTD_Model_VEA, SinPhi, SinTheta, V, adjointBornAccumulation, coherence,
    device_ptr, forwardBornInjection, free, propagateadjoint, propagateforward, propagateforward_linear,
    receiver_data, scale_spatial_derivatives, set_fused_step, set_fused_tile, set_receivers, set_receivers_interp, set_sources, set_sources_interp, size,
    sync_to_device, sync_to_host, synchronize, timestep_inject_extract, timestep_inject_extract_graph)
from .spacetime import (  # noqa: F401
    SourcePoint32, causal_ricker, extractdata, hickscoeffs, injectdata, interpadjoint, interpfilters,
    interpforward, linearcoeffs, receiver_blocking, ricker, source_blocking)
from .tuning import autotune, fused_config, set_fused_config, tilesize  # noqa: F401,E402
