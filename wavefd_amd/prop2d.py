"""The 2D propagators' host API, mirroring the reference's Julia wrappers
(src/prop2DAco{Iso,VTI,TTI}DenQ_DEO2_FDTD.jl) over the drop-in libraries
libprop2DAco{Iso,VTI,TTI}DenQ_DEO2_FDTD.

Arrays are (nz, nx) in Julia's column-major order: numpy views of shape (nz, nx) with Fortran
order over the library's host mirrors (k = kx nz + kz).  The default host coherence gives the
reference's synchronous semantics: every call uploads the host arrays it reads and downloads
the ones it writes.
"""
import ctypes
import os

import numpy as np

from . import _lib
from .prop3d import (ImagingConditionStandard, ImagingConditionWaveFieldSeparationFWI,  # noqa: F401
                     ImagingConditionWaveFieldSeparationMIX, ImagingConditionWaveFieldSeparationRTM, _ptr)

_EPS, _ETA = "ϵ", "η"  # the reference's dictionary keys: ϵ and η


class _Prop2D:
    """src/prop2DAco{Iso,VTI,TTI}DenQ_DEO2_FDTD.jl:5-42 (the constructors are identical)"""
    KIND = None

    def __init__(self, *, nz=0, nx=0, nsponge=60, nbz=512, nbx=8, dz=5.0, dx=5.0, dt=0.001, nthreads=None,
                 freesurface=True, freqQ=5.0, qMin=0.1, qInterior=100.0):
        if nthreads is None:
            nthreads = os.cpu_count() or 1
        nz, nx, nsponge, nbz, nbx, nthreads = (int(round(v)) for v in (nz, nx, nsponge, nbz, nbx, nthreads))
        dz, dx, dt = (np.float32(v) for v in (dz, dx, dt))
        assert nx > 0
        assert nz > 0
        assert nsponge > 0
        assert nthreads > 0
        assert nbx > 0
        assert nbz > 0
        fs = 1 if freesurface else 0
        self._lib = _lib.load(self.KIND)
        self.p = self._lib.alloc(fs, nthreads, nx, nz, nsponge, dx, dz, dt, nbx, nbz)
        if not self.p:
            raise RuntimeError("alloc failed: " + _lib.last_error(self.KIND))
        self._lib.ClearError()
        self._lib.SetupDtOmegaInvQ(self.p, freqQ, qMin, qInterior)
        err = _lib.last_error(self.KIND)
        if err:  # freqQ < FLT_EPSILON (src/absorb.h:98-103): the reference exits; this raises
            self._lib.ClearError()
            self._lib.free(self.p)
            self.p = None
            raise ValueError(err.strip())
        self.shape = (nz, nx)
        self.n = nz * nx

    def __del__(self):
        try:
            free(self)
        except Exception:
            pass

    def __repr__(self):
        nz, nx = self.shape
        return "%s -- nx,nz; %d,%d" % (type(self).__name__, nx, nz)

    def _check(self):
        msg = _lib.last_error(self.KIND)
        if msg:
            self._lib.ClearError()
            raise RuntimeError(msg)


class Prop2DAcoIsoDenQ_DEO2_FDTD(_Prop2D):
    """src/prop2DAcoIsoDenQ_DEO2_FDTD.jl:1-41"""
    KIND = "iso2d"


class Prop2DAcoVTIDenQ_DEO2_FDTD(_Prop2D):
    """src/prop2DAcoVTIDenQ_DEO2_FDTD.jl:1-42"""
    KIND = "vti2d"


class Prop2DAcoTTIDenQ_DEO2_FDTD(_Prop2D):
    """src/prop2DAcoTTIDenQ_DEO2_FDTD.jl:1-42"""
    KIND = "tti2d"


class Prop2DAcoIsoDenQ_DEO2_FDTD_Model_V:
    pass


class Prop2DAcoIsoDenQ_DEO2_FDTD_Model_B:
    pass


class Prop2DAcoIsoDenQ_DEO2_FDTD_Model_VB:
    pass


class Prop2DAcoVTIDenQ_DEO2_FDTD_Model_V:
    pass


class Prop2DAcoVTIDenQ_DEO2_FDTD_Model_VEA:
    pass


class Prop2DAcoTTIDenQ_DEO2_FDTD_Model_V:
    pass


class Prop2DAcoTTIDenQ_DEO2_FDTD_Model_VEA:
    pass


def free(prop):
    if getattr(prop, "p", None):
        prop._lib.free(prop.p)
        prop.p = None


def size(prop):
    """(nz, nx), src/prop2DAcoIsoDenQ_DEO2_FDTD.jl:45-49"""
    return (prop._lib.getNz(prop.p), prop._lib.getNx(prop.p))


def _wrap(prop, getter):
    fn = getattr(prop._lib, getter, None)
    if fn is None:
        raise AttributeError("%s has no %s" % (type(prop).__name__, getter))
    addr = fn(prop.p)
    if not addr:
        prop._check()
        raise RuntimeError("getter %s returned NULL" % getter)
    buf = (ctypes.c_float * prop.n).from_address(addr)
    return np.ndarray(prop.shape, dtype=np.float32, buffer=buf, order="F")


def _make_getter(name):
    def getter(prop):
        return _wrap(prop, "get" + name)

    getter.__name__ = name
    getter.__doc__ = "Host view of the library-owned %s array, shape (nz, nx), Fortran order." % name
    return getter


V = _make_getter("V")
B = _make_getter("B")
DtOmegaInvQ = _make_getter("DtOmegaInvQ")
PSpace = _make_getter("PSpace")
PCur = _make_getter("PCur")
POld = _make_getter("POld")
# VTI and TTI (src/prop2DAcoVTIDenQ_DEO2_FDTD.jl:52-55, src/prop2DAcoTTIDenQ_DEO2_FDTD.jl:52-55)
Eps = _make_getter("Eps")
Eta = _make_getter("Eta")
F = _make_getter("F")
MSpace = _make_getter("MSpace")
MCur = _make_getter("MCur")
MOld = _make_getter("MOld")
SinTheta = _make_getter("SinTheta")
CosTheta = _make_getter("CosTheta")


def propagateforward(prop):
    prop._lib.TimeStep(prop.p)
    prop._check()


def propagateforward_linear(prop):
    """VTI only (src/prop2DAcoVTIDenQ_DEO2_FDTD.jl:58)"""
    if prop.KIND != "vti2d":
        raise TypeError("%s has no linear step" % type(prop).__name__)
    prop._lib.TimeStepLinear(prop.p)
    prop._check()


def propagateadjoint(prop):
    """self-adjoint: the linear step for VTI (src/prop2DAcoVTIDenQ_DEO2_FDTD.jl:59), the step
    otherwise (src/prop2DAcoIsoDenQ_DEO2_FDTD.jl:61, src/prop2DAcoTTIDenQ_DEO2_FDTD.jl:58)"""
    if prop.KIND == "vti2d":
        propagateforward_linear(prop)
    else:
        propagateforward(prop)


def scale_spatial_derivatives(prop):
    prop._lib.ScaleSpatialDerivatives(prop.p)
    prop._check()


def forwardBornInjection(prop, modeltype, dmodel, wavefields):
    """src/prop2DAcoIsoDenQ_DEO2_FDTD.jl:66-85, src/prop2DAcoVTIDenQ_DEO2_FDTD.jl:68-87 (TTI the same)"""
    L, p, n = prop._lib, prop.p, prop.n
    name = type(modeltype).__name__
    if prop.KIND == "iso2d":
        if name.endswith("_Model_V"):
            L.ForwardBornInjection_V(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n))
        elif name.endswith("_Model_VB"):
            L.ForwardBornInjection_VB(p, _ptr(dmodel["v"], n), _ptr(dmodel["b"], n), _ptr(wavefields["pold"], n),
                                      _ptr(wavefields["pspace"], n))
        elif name.endswith("_Model_B"):
            L.ForwardBornInjection_B(p, _ptr(dmodel["b"], n), _ptr(wavefields["pold"], n),
                                     _ptr(wavefields["pspace"], n))
        else:
            raise TypeError(name)
    elif name.endswith("_Model_V"):
        L.ForwardBornInjection_V(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n), _ptr(wavefields["mspace"], n))
    elif name.endswith("_Model_VEA"):
        L.ForwardBornInjection_VEA(p, _ptr(dmodel["v"], n), _ptr(dmodel[_EPS], n), _ptr(dmodel[_ETA], n),
                                   _ptr(wavefields["pold"], n), _ptr(wavefields["mold"], n),
                                   _ptr(wavefields["pspace"], n), _ptr(wavefields["mspace"], n))
    else:
        raise TypeError(name)
    prop._check()


def adjointBornAccumulation(prop, modeltype, imagingcondition, dmodel, wavefields):
    """src/prop2DAcoIsoDenQ_DEO2_FDTD.jl:87-144, src/prop2DAcoVTIDenQ_DEO2_FDTD.jl:75-122 (TTI the same)"""
    L, p, n = prop._lib, prop.p, prop.n
    name = type(modeltype).__name__
    ic = type(imagingcondition).__name__
    if ic == "ImagingConditionWaveFieldSeparationMIX":
        keys = {"v": ("all_v", "rtm_v"), "b": ("all_b", "rtm_b"), _EPS: ("all_" + _EPS, "rtm_" + _EPS),
                _ETA: ("all_" + _ETA, "rtm_" + _ETA)}
        d_all = {k: dmodel[a] for k, (a, r) in keys.items() if a in dmodel}
        d_rtm = {k: dmodel[r] for k, (a, r) in keys.items() if r in dmodel}
        adjointBornAccumulation(prop, modeltype, ImagingConditionStandard(), d_all, wavefields)
        adjointBornAccumulation(prop, modeltype, ImagingConditionWaveFieldSeparationRTM(), d_rtm, wavefields)
        return
    if ic in ("ImagingConditionWaveFieldSeparationFWI", "ImagingConditionWaveFieldSeparationRTM"):
        if not name.endswith("_Model_V"):
            raise TypeError("wavefield separation images v only")
        isfwi = 1 if ic.endswith("FWI") else 0
        if prop.KIND == "iso2d":
            L.AdjointBornAccumulation_wavefieldsep(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n), isfwi)
        else:
            L.AdjointBornAccumulation_wavefieldsep_V(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n),
                                                     _ptr(wavefields["mspace"], n), isfwi)
    elif prop.KIND == "iso2d":
        if name.endswith("_Model_V"):
            L.AdjointBornAccumulation_V(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n))
        elif name.endswith("_Model_VB"):
            L.AdjointBornAccumulation_VB(p, _ptr(dmodel["v"], n), _ptr(dmodel["b"], n), _ptr(wavefields["pold"], n),
                                         _ptr(wavefields["pspace"], n))
        elif name.endswith("_Model_B"):
            L.AdjointBornAccumulation_B(p, _ptr(dmodel["b"], n), _ptr(wavefields["pold"], n),
                                        _ptr(wavefields["pspace"], n))
        else:
            raise TypeError(name)
    elif name.endswith("_Model_V"):
        L.AdjointBornAccumulation_V(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n),
                                    _ptr(wavefields["mspace"], n))
    elif name.endswith("_Model_VEA"):
        L.AdjointBornAccumulation_VEA(p, _ptr(dmodel["v"], n), _ptr(dmodel[_EPS], n), _ptr(dmodel[_ETA], n),
                                      _ptr(wavefields["pold"], n), _ptr(wavefields["mold"], n),
                                      _ptr(wavefields["pspace"], n), _ptr(wavefields["mspace"], n))
    else:
        raise TypeError(name)
    prop._check()


def applyFirstDerivatives2D_PlusHalf(freesurface, nx, nz, c8, invDx, invDz, inX, inZ, outX, outZ, nthreads=1,
                                     nbx=8, nbz=512):
    """The stateless D+ on (nz, nx) host arrays with caller coefficients c8 = (c1, c2, c3, c4)
    (src/prop2DAcoVTIDenQ_DEO2_FDTD.cc:158-176)."""
    L = _lib.load("vti2d")
    n = nx * nz
    L.applyFirstDerivatives2D_PlusHalf(1 if freesurface else 0, nx, nz, nthreads, *[np.float32(c) for c in c8],
                                       np.float32(invDx), np.float32(invDz), _ptr(inX, n), _ptr(inZ, n),
                                       _ptr(outX, n), _ptr(outZ, n), nbx, nbz)
    msg = _lib.last_error("vti2d")
    if msg:
        L.ClearError()
        raise RuntimeError(msg)


# ---- host coherence / device path (additive) ------------------------------------------

def coherence(prop, mode):
    """Set the host-coherence mode: 'coherent' (default, the reference's synchronous host
    arrays) or 'device' (calls stream-ordered on the handle's stream, host arrays untouched)."""
    prop._lib.SetHostCoherence(prop.p, 1 if mode == "coherent" else 0)


def sync_to_device(prop):
    prop._lib.SyncToDevice(prop.p)
    prop._check()


def sync_to_host(prop):
    prop._lib.SyncToHost(prop.p)
    prop._check()


def synchronize(prop):
    prop._lib.Synchronize(prop.p)
    prop._check()


def stream(prop):
    """The handle's HIP stream (an integer address)."""
    return prop._lib.GetStream(prop.p)


def device_ptr(prop, name):
    """Device address of a state array by name (v b dtOmegaInvQ eps eta f sinTheta cosTheta
    pCur pOld mCur mOld pSpace mSpace)."""
    addr = prop._lib.GetDevicePtr(prop.p, name.encode())
    if not addr:
        raise KeyError(name)
    return addr


def set_fused_step(prop, fused=True):
    """Choose the fused single-pass 2D step (default) or the two passes; the results are the same."""
    prop._lib.SetFusedStep(prop.p, 1 if fused else 0)
