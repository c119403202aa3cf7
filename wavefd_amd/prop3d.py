"""Host API of the 3D propagators, mirroring the reference's Julia wrappers.

Same constructor keywords and defaults, accessors, step and Born entry points as
src/prop3DAco{Iso,VTI,TTI}DenQ_DEO2_FDTD.jl, calling the gfx950 libraries through the
reference's C-ABI.  Julia's ``!`` suffixes are dropped (``propagateforward``), and the
accessors (``V(prop)``, ``PCur(prop)``, ...) return numpy views of the library-owned host
arrays with Julia's shape and order: ``(nz, ny, nx)`` in column-major (Fortran) order, so
``V(prop)[iz, iy, ix]`` addresses the same cell as in Julia (0-based).

Host coherence: by default every call keeps the host arrays and the device state in step,
as the reference's single address space does (`coherence(prop, "coherent")`).  For the
fast path, ``coherence(prop, "device")`` makes calls stream-ordered and leaves host arrays
alone until ``sync_to_host(prop)`` / ``sync_to_device(prop)``; ``device_ptr(prop, name)``
returns the device address of a state array.
"""
import ctypes
import os

import numpy as np

from . import _lib


class ImagingConditionStandard:
    pass


class ImagingConditionWaveFieldSeparationFWI:
    pass


class ImagingConditionWaveFieldSeparationRTM:
    pass


class ImagingConditionWaveFieldSeparationMIX:
    pass


def _ptr(a, n=None):
    """float* address of a numpy float32 array, a torch tensor (host or device) or an int."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):  # torch.Tensor
        return int(a.data_ptr())
    if isinstance(a, np.ndarray):
        if a.dtype != np.float32:
            raise TypeError("expected float32 array, got %s" % a.dtype)
        if not (a.flags.c_contiguous or a.flags.f_contiguous):
            raise ValueError("array must be contiguous")
        if n is not None and a.size != n:
            raise ValueError("array has %d elements, expected %d" % (a.size, n))
        return a.ctypes.data
    raise TypeError("unsupported array type %r" % type(a))


class _Prop3D:
    KIND = None
    ARRAYS = ()

    def __init__(self, *, nz=0, ny=0, nx=0, nsponge=60, nbz=512, nby=8, nbx=8, dz=5.0, dy=8.0, dx=5.0,
                 dt=0.001, nthreads=None, freesurface=True, freqQ=5.0, qMin=0.1, qInterior=100.0):
        # src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:5-45
        if nthreads is None:
            nthreads = os.cpu_count() or 1
        nz, ny, nx, nsponge, nbz, nby, nbx, nthreads = (int(round(v)) for v in
                                                         (nz, ny, nx, nsponge, nbz, nby, nbx, nthreads))
        dz, dy, dx, dt = (np.float32(v) for v in (dz, dy, dx, dt))
        assert nx > 0
        assert ny > 0
        assert nz > 0
        assert nsponge > 0
        assert nthreads > 0
        assert nbx > 0
        assert nby > 0
        assert nbz > 0
        fs = 1 if freesurface else 0
        self._lib = _lib.load(self.KIND)
        self.p = self._lib.alloc(fs, nthreads, nx, ny, nz, nsponge, dx, dy, dz, dt, nbx, nby, nbz)
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
        self.shape = (nz, ny, nx)
        self.n = nz * ny * nx
        self.dt = dt

    def __del__(self):
        try:
            free(self)
        except Exception:
            pass

    def __repr__(self):
        nz, ny, nx = self.shape
        return "%s -- nx,ny,nz; %d,%d,%d" % (type(self).__name__, nx, ny, nz)

    def _check(self):
        msg = _lib.last_error(self.KIND)
        if msg:
            self._lib.ClearError()
            raise RuntimeError(msg)


def free(prop):
    if getattr(prop, "p", None):
        prop._lib.free(prop.p)
        prop.p = None


def size(prop):
    """(nz, ny, nx), src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:49-54"""
    lib = prop._lib
    return (lib.getNz(prop.p), lib.getNy(prop.p), lib.getNx(prop.p))


def _wrap(prop, getter):
    addr = getattr(prop._lib, getter)(prop.p)
    if not addr:
        prop._check()
        raise RuntimeError("getter %s returned NULL" % getter)
    nz, ny, nx = prop.shape
    buf = (ctypes.c_float * prop.n).from_address(addr)
    return np.ndarray((nz, ny, nx), dtype=np.float32, buffer=buf, order="F")


def _make_getter(name):
    def getter(prop):
        return _wrap(prop, "get" + name)

    getter.__name__ = name
    getter.__doc__ = "Host view of the library-owned %s array, shape (nz, ny, nx), Fortran order." % name
    return getter


V = _make_getter("V")
B = _make_getter("B")
DtOmegaInvQ = _make_getter("DtOmegaInvQ")
PSpace = _make_getter("PSpace")
PCur = _make_getter("PCur")
POld = _make_getter("POld")
Eps = _make_getter("Eps")
Eta = _make_getter("Eta")
F = _make_getter("F")
MSpace = _make_getter("MSpace")
MCur = _make_getter("MCur")
MOld = _make_getter("MOld")
SinTheta = _make_getter("SinTheta")
CosTheta = _make_getter("CosTheta")
SinPhi = _make_getter("SinPhi")
CosPhi = _make_getter("CosPhi")


def propagateforward(prop):
    prop._lib.TimeStep(prop.p)
    prop._check()


def propagateforward_linear(prop):
    """VTI only (src/prop3DAcoVTIDenQ_DEO2_FDTD.jl:63)."""
    prop._lib.TimeStepLinear(prop.p)
    prop._check()


def propagateadjoint(prop):
    # Iso/TTI: self-adjoint (src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:69); VTI: the linear step
    # (src/prop3DAcoVTIDenQ_DEO2_FDTD.jl:65)
    if isinstance(prop, Prop3DAcoVTIDenQ_DEO2_FDTD):
        propagateforward_linear(prop)
    else:
        propagateforward(prop)


def scale_spatial_derivatives(prop):
    prop._lib.ScaleSpatialDerivatives(prop.p)
    prop._check()


# ---- host coherence / device path (additive) ------------------------------------------

def coherence(prop, mode=None):
    """Get or set the host-coherence mode: 'coherent' (default) or 'device'."""
    if mode is not None:
        prop._lib.SetHostCoherence(prop.p, 1 if mode == "coherent" else 0)
    return "coherent" if prop._lib.GetHostCoherence(prop.p) else "device"


def sync_to_device(prop):
    prop._lib.SyncToDevice(prop.p)
    prop._check()


def sync_to_host(prop):
    prop._lib.SyncToHost(prop.p)
    prop._check()


def synchronize(prop):
    prop._lib.Synchronize(prop.p)
    prop._check()


def device_ptr(prop, name):
    """Device address of a state array: v b dtOmegaInvQ eps eta f sinTheta cosTheta sinPhi
    cosPhi pCur pOld mCur mOld pSpace mSpace (after swaps, pCur is the newest field)."""
    addr = prop._lib.GetDevicePtr(prop.p, name.encode())
    if not addr:
        raise KeyError(name)
    return addr


def set_fused_step(prop, fused=True):
    """Choose the fused single-pass step (default) or the two-pass operator step."""
    prop._lib.SetFusedStep(prop.p, 1 if fused else 0)


FUSED_TILES = {"auto": 0, "64x16": 1, "32x32": 2, "wide": 3, "wide2": 4}


def set_fused_tile(prop, tile="auto"):
    """Tile shape of the fused step kernel: "64x16" (64 z-lanes x 16 y-rows per workgroup,
    57 x 9 outputs), "32x32" (25 x 25 outputs), "wide" (512 threads with 4 y-rows each for Iso,
    57 x 25 outputs; 3 for VTI and TTI, 57 x 17), "wide2" (Iso's one-wavefield steps: 2 y-rows
    per thread, 57 x 9 outputs, two workgroups per CU; other steps take "wide") or "auto".
    Results do not depend on it."""
    if prop._lib.SetFusedTile(prop.p, FUSED_TILES[tile]) != 0:
        prop._check()


def _partitions_args(partitions):
    from .spacetime import _partition_arrays
    return _partition_arrays(partitions)


def set_sources(prop, partitions, data, fields=1):
    """Register injection points (partitions from source_blocking) and their data
    ((nt, nr) Julia layout == (nr, nt) C-order float32) for timestep_inject_extract."""
    keep, parts_pp, nparts, npts = _partitions_args(partitions)
    data = np.ascontiguousarray(data, np.float32) if isinstance(data, np.ndarray) else data
    nt = data.shape[-1] if isinstance(data, np.ndarray) and data.ndim == 2 else None
    if nt is None:
        raise ValueError("data must be a 2D (nr, nt) C-ordered float32 array")
    rc = prop._lib.SetSources(prop.p, parts_pp, nparts, npts, _ptr(data), nt, fields)
    if rc != 0:
        prop._check()
    prop._src_keep = keep


def set_receivers(prop, partitions, nt, fields=1):
    keep, parts_pp, nparts, npts = _partitions_args(partitions)
    rc = prop._lib.SetReceivers(prop.p, parts_pp, nparts, npts, nt, fields)
    if rc != 0:
        prop._check()
    prop._rcv_keep = keep
    prop._rcv_shape = None


def set_sources_interp(prop, partitions, data, h, nm, fields=1):
    """Register injection points whose data ((nd, nr) Julia layout == (nr, nd) C-order float32)
    is at the coarse rate: each injection at fine step it adds the reference's interpadjoint of
    the data at it (filters h = interpfilters(dtmod, dtrec), nm fine samples)."""
    keep, parts_pp, nparts, npts = _partitions_args(partitions)
    data = np.ascontiguousarray(data, np.float32)
    if data.ndim != 2:
        raise ValueError("data must be a 2D (nr, nd) C-ordered float32 array")
    h = np.ascontiguousarray(h, np.float32)
    rc = prop._lib.SetSourcesInterp(prop.p, parts_pp, nparts, npts, data.ctypes.data, data.shape[1], h.ctypes.data,
                                    h.shape[0], nm, fields)
    if rc != 0:
        prop._check()
    prop._src_keep = keep


def set_receivers_interp(prop, partitions, nd, h, nm, fields=1):
    """Register extraction points with a coarse (nr, nd) trace buffer: each extraction at fine
    step it adds the reference's interpforward contributions of that fine sample."""
    keep, parts_pp, nparts, npts = _partitions_args(partitions)
    h = np.ascontiguousarray(h, np.float32)
    rc = prop._lib.SetReceiversInterp(prop.p, parts_pp, nparts, npts, nd, h.ctypes.data, h.shape[0], nm, fields)
    if rc != 0:
        prop._check()
    prop._rcv_keep = keep
    prop._rcv_shape = None


def timestep_inject_extract(prop, it):
    """propagateforward, then inject the registered sources at it (1-based) into the new
    pcur, then extract the registered receivers at it."""
    prop._lib.TimeStepInjectExtract(prop.p, it)
    prop._check()


def timestep_inject_extract_graph(prop, it0, nsteps):
    """nsteps of timestep_inject_extract (time indices it0 .. it0+nsteps-1, 1-based) captured
    as one HIP graph and launched once (DEVICE mode, fused step); the same results as the
    calls one by one, with one host launch instead of three per step."""
    if prop._lib.TimeStepInjectExtractGraph(prop.p, int(it0), int(nsteps)) != 0:
        prop._check()
        raise RuntimeError("TimeStepInjectExtractGraph failed")


def receiver_data(prop, nr, nt):
    out = np.zeros((nr, nt), np.float32)
    if prop._lib.GetReceiverData(prop.p, out.ctypes.data) != 0:
        prop._check()
    return out


# ---- the three propagators ---------------------------------------------------------------

class Prop3DAcoIsoDenQ_DEO2_FDTD(_Prop3D):
    """src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:1-47"""
    KIND = "iso"


class Prop3DAcoVTIDenQ_DEO2_FDTD(_Prop3D):
    """src/prop3DAcoVTIDenQ_DEO2_FDTD.jl:1-47"""
    KIND = "vti"


class Prop3DAcoTTIDenQ_DEO2_FDTD(_Prop3D):
    """src/prop3DAcoTTIDenQ_DEO2_FDTD.jl:1-47"""
    KIND = "tti"


class Prop3DAcoIsoDenQ_DEO2_FDTD_Model_V:
    pass


class Prop3DAcoIsoDenQ_DEO2_FDTD_Model_B:
    pass


class Prop3DAcoIsoDenQ_DEO2_FDTD_Model_VB:
    pass


class Prop3DAcoVTIDenQ_DEO2_FDTD_Model_V:
    pass


class Prop3DAcoVTIDenQ_DEO2_FDTD_Model_VEA:
    pass


class Prop3DAcoTTIDenQ_DEO2_FDTD_Model_V:
    pass


class Prop3DAcoTTIDenQ_DEO2_FDTD_Model_VEA:
    pass


_EPS, _ETA = "ϵ", "η"  # the reference's dictionary keys: ϵ and η


def forwardBornInjection(prop, modeltype, dmodel, wavefields):
    """src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:75-94, src/prop3DAcoVTIDenQ_DEO2_FDTD.jl:75-93,
    src/prop3DAcoTTIDenQ_DEO2_FDTD.jl (same shapes)."""
    L, p, n = prop._lib, prop.p, prop.n
    name = type(modeltype).__name__
    if name.endswith("_Model_V") and prop.KIND == "iso":
        L.ForwardBornInjection_V(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n))
    elif name.endswith("_Model_VB"):
        L.ForwardBornInjection_VB(p, _ptr(dmodel["v"], n), _ptr(dmodel["b"], n), _ptr(wavefields["pold"], n),
                                  _ptr(wavefields["pspace"], n))
    elif name.endswith("_Model_B"):
        L.ForwardBornInjection_B(p, _ptr(dmodel["b"], n), _ptr(wavefields["pold"], n), _ptr(wavefields["pspace"], n))
    elif name.endswith("_Model_V"):
        L.ForwardBornInjection_V(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n),
                                 _ptr(wavefields["mspace"], n))
    elif name.endswith("_Model_VEA"):
        L.ForwardBornInjection_VEA(p, _ptr(dmodel["v"], n), _ptr(dmodel[_EPS], n), _ptr(dmodel[_ETA], n),
                                   _ptr(wavefields["pold"], n), _ptr(wavefields["mold"], n),
                                   _ptr(wavefields["pspace"], n), _ptr(wavefields["mspace"], n))
    else:
        raise TypeError(name)
    prop._check()


def adjointBornAccumulation(prop, modeltype, imagingcondition, dmodel, wavefields):
    """src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:96-150 and the VTI/TTI twins."""
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
        isfwi = 1 if ic.endswith("FWI") else 0
        if prop.KIND == "iso":
            L.AdjointBornAccumulation_wavefieldsep(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n), isfwi)
        else:
            L.AdjointBornAccumulation_wavefieldsep_V(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n),
                                                     _ptr(wavefields["mspace"], n), isfwi)
        prop._check()
        return
    if prop.KIND == "iso":
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
    else:
        if name.endswith("_Model_V"):
            L.AdjointBornAccumulation_V(p, _ptr(dmodel["v"], n), _ptr(wavefields["pspace"], n),
                                        _ptr(wavefields["mspace"], n))
        elif name.endswith("_Model_VEA"):
            L.AdjointBornAccumulation_VEA(p, _ptr(dmodel["v"], n), _ptr(dmodel[_EPS], n), _ptr(dmodel[_ETA], n),
                                          _ptr(wavefields["pold"], n), _ptr(wavefields["mold"], n),
                                          _ptr(wavefields["pspace"], n), _ptr(wavefields["mspace"], n))
        else:
            raise TypeError(name)
    prop._check()
