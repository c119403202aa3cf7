"""Born modelling and imaging of one shot on a ginsu window, on the device.

The reference ships the single-step pieces (the propagators, ``forwardBornInjection_V`` and
``adjointBornAccumulation_V``, src/prop3DAcoIsoDenQ_DEO2_FDTD.h:305-334 and :518-546, and their
VTI twins) and leaves the time loops to its callers (JetPackWaveFD, out of tree).  This driver
writes those loops against this package's C-ABI, all state in HBM:

1. background forward: per step ``TimeStep``, inject the source, ``ScaleSpatialDerivatives``;
   DP_n = v^2/b L(p_{n-1}) (and DM_n) saved to a device snapshot buffer (by default through the
   fused ``TimeStepSaveDP``, and likewise the fused pairs of passes 2 and 3);
2. Born forward (the linearisation J in dv): per step ``TimeStep``, ``ForwardBornInjection_V``
   with DP_n (p += 2 dt^2 dv/v DP_n into the new field), extract the receivers: d = J dv;
3. adjoint (J^T): the forward step is self-adjoint in the b/(dt^2 v^2)-weighted inner product,
   so J^T d runs the same step on the time-reversed data injected at the receivers with
   weights dt^2 v^2 / b, and images with ``AdjointBornAccumulation_V`` (g += 2b/v^3 DP_n lam_n,
   which carries the weight b/(dt^2 v^2) of the sources' side).  That call images the field
   in pOld, i.e. the previous step's field, so DP_{nt+2-m} is used after reversed step m,
   and one extra step images DP_1;
4. the window's image is added into the shared image over the window's interior
   (``Window.interior``).

``tests/test_gpu_born_pipeline.py`` checks <J dv, d> = <dv, J^T d> and J against a central
difference of the full modelling.
"""
import time

import numpy as np

from . import ginsu, prop3d, spacetime

F32 = np.float32
_CLS = {"iso": prop3d.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": prop3d.Prop3DAcoVTIDenQ_DEO2_FDTD}
MODEL_FIELDS = {"iso": ("v", "b"), "vti": ("v", "b", "eps", "eta", "f")}


def _shift(points_z, points_y, points_x, window, spacing):
    oz, oy, ox = window.origin
    dz, dy, dx = spacing
    return (np.asarray(points_z, np.float64) - oz * dz, np.asarray(points_y, np.float64) - oy * dy,
            np.asarray(points_x, np.float64) - ox * dx)


def _gather(field_dev, pts, nsteps=1):
    """values of a device field at the taps of pts (one per tap, in tap order), via the
    device extraction path of libspacetime"""
    taps = pts.copy()
    taps["ir"] = np.arange(1, len(taps) + 1)
    taps["c"] = 1.0
    buf = ginsu.DeviceBuffer(max(4, 4 * len(taps)))
    buf.zero()
    buf.sync()
    spacetime.extractdata(_DevArray(buf.ptr, (len(taps), 1)), field_dev, 1, [taps])
    out = np.zeros(len(taps), F32)
    buf.download(out)
    buf.free()
    return out


class _DevArray:
    """a device address with the (nr, nt) shape that libspacetime's wrappers read nt from"""

    def __init__(self, ptr, shape):
        self.ptr, self.shape = ptr, shape

    def data_ptr(self):
        return self.ptr


class BornShot:
    """One shot: source (z, y, x), receiver coordinate arrays, a wavelet of nt samples at dt, all
    in the big grid's coordinates (origin 0)."""

    def __init__(self, src, rcv_z, rcv_y, rcv_x, wavelet):
        self.src = tuple(float(c) for c in src)
        self.rz, self.ry, self.rx = (np.asarray(a, np.float64) for a in (rcv_z, rcv_y, rcv_x))
        self.wavelet = np.ascontiguousarray(np.asarray(wavelet, F32).reshape(1, -1))

    @property
    def nr(self):
        return len(self.rz)


def born_shot(physics, window, spacing, dt, shot, model, dmodel, image, nsponge=20, freesurface=True, freqQ=5.0,
              qMin=0.1, qInterior=100.0, snapshots=None, adjoint_data=None, keep_born_data=True, fused=True):
    """Run the three passes of one shot on `window` and add its image into `image`.

    model: {'v','b'[,'eps','eta','f']} -> device address of the big grid's field; dmodel: device
    address of the big dv; image: device address of the big gradient (accumulated, +=).
    snapshots: optional DeviceBuffer of at least nt * window cells * 4 B (x2 for VTI), reused
    across shots.  adjoint_data: (nr, nt) float32 data for the adjoint pass (default: the Born
    data of pass 2, i.e. image = J^T J dv).  fused: each pass's per-step pair as one fused step
    (TimeStepSaveDP, TimeStepBornInjection_V, TimeStepAdjointBornAccumulation_V; the same
    results as the separate calls).  Returns a dict with the Born data and timings."""
    if physics not in _CLS:
        raise ValueError("physics must be 'iso' or 'vti'")
    nf = 1 if physics == "iso" else 2
    nz, ny, nx = window.size
    dz, dy, dx = spacing
    nt = shot.wavelet.shape[1]
    ncell = window.ncells
    t0 = time.perf_counter()
    prop = _CLS[physics](nz=nz, ny=ny, nx=nx, nsponge=nsponge, dz=dz, dy=dy, dx=dx, dt=dt, freesurface=freesurface,
                         freqQ=freqQ, qMin=qMin, qInterior=qInterior)
    prop3d.coherence(prop, "device")
    stream = prop._lib.GetStream(prop.p)
    for name in MODEL_FIELDS[physics]:
        ginsu.extract(model[name], window, prop3d.device_ptr(prop, name), stream)
    own_snap = snapshots is None
    if own_snap:
        snapshots = ginsu.DeviceBuffer(4 * ncell * nt * nf)
    if snapshots.nbytes < 4 * ncell * nt * nf:
        raise ValueError("snapshot buffer too small")
    dv = ginsu.DeviceBuffer(4 * ncell)
    ginsu.extract(dmodel, window, dv.ptr, stream)
    gimg = ginsu.DeviceBuffer(4 * ncell)
    gimg.zero(stream)
    wf = ("pCur", "pOld") + (("mCur", "mOld") if nf == 2 else ())

    def snap(kind, n):  # address of DP_n / DM_n, n 1-based
        return snapshots.ptr + 4 * ncell * ((n - 1) * nf + kind)

    def zero_fields():
        for name in wf:
            ginsu.zero(prop3d.device_ptr(prop, name), 4 * ncell, stream)

    # receivers and source taps in window coordinates
    sz, sy, sx = _shift([shot.src[0]], [shot.src[1]], [shot.src[2]], window, spacing)
    fsi = 1 if freesurface else 0  # 1-based free-surface row; 0: no image terms
    src_pts = spacetime.hickscoeffs(F32(dz), F32(dy), F32(dx), 0.0, 0.0, 0.0, nz, ny, nx, sz, sy, sx, fs_index=fsi)
    rz, ry, rx = _shift(shot.rz, shot.ry, shot.rx, window, spacing)
    rcv_pts = spacetime.linearcoeffs(F32(dz), F32(dy), F32(dx), 0.0, 0.0, 0.0, nz, ny, nx, rz, ry, rx,
                                     fs_index=fsi)
    lib = prop._lib
    # 1. background forward, saving DP (and DM)
    prop3d.set_sources(prop, [src_pts], shot.wavelet, fields=1 if nf == 1 else 3)
    prop3d.set_receivers(prop, [], nt)
    t1 = time.perf_counter()
    for it in range(1, nt + 1):
        if fused:
            if nf == 1:
                lib.TimeStepSaveDP(prop.p, snap(0, it))
            else:
                lib.TimeStepSaveDP(prop.p, snap(0, it), snap(1, it))
            lib.InjectRegistered(prop.p, it)
        else:
            lib.TimeStepInjectExtract(prop.p, it)
            lib.ScaleSpatialDerivatives(prop.p)
            ginsu.copy(snap(0, it), prop3d.device_ptr(prop, "pSpace"), 4 * ncell, stream)
            if nf == 2:
                ginsu.copy(snap(1, it), prop3d.device_ptr(prop, "mSpace"), 4 * ncell, stream)
    prop._check()
    # 2. Born forward: d = J dv
    zero_fields()
    prop3d.set_sources(prop, [], np.zeros((1, nt), F32))
    prop3d.set_receivers(prop, [rcv_pts], nt)
    t2 = time.perf_counter()
    for it in range(1, nt + 1):
        dps = (snap(0, it),) if nf == 1 else (snap(0, it), snap(1, it))
        if fused:
            lib.TimeStepBornInjection_V(prop.p, dv.ptr, *dps)
        else:
            lib.TimeStep(prop.p)
            lib.ForwardBornInjection_V(prop.p, dv.ptr, *dps)
        lib.ExtractRegistered(prop.p, it)
    prop._check()
    born_data = prop3d.receiver_data(prop, shot.nr, nt) if (keep_born_data or adjoint_data is None) else None
    # 3. adjoint with imaging: receivers inject with weights dt^2 v^2 / b, time reversed
    d = born_data if adjoint_data is None else np.ascontiguousarray(adjoint_data, F32)
    prop3d.synchronize(prop)  # the gather below runs on the null stream
    vb = {k: _gather(prop3d.device_ptr(prop, k), rcv_pts) for k in ("v", "b")}
    adj_pts = rcv_pts.copy()
    adj_pts["c"] = (rcv_pts["c"].astype(np.float64) * dt * dt * vb["v"].astype(np.float64) ** 2 /
                    vb["b"].astype(np.float64)).astype(F32)
    zero_fields()
    prop3d.set_receivers(prop, [], nt)
    # R^T of the P-field extraction: the adjoint source goes into P only
    prop3d.set_sources(prop, [adj_pts], np.ascontiguousarray(d[:, ::-1]), fields=1)
    t3 = time.perf_counter()
    for m in range(1, nt + 2):
        n = nt + 2 - m
        dps = () if m == 1 else ((snap(0, n),) if nf == 1 else (snap(0, n), snap(1, n)))
        if m >= 2 and fused:  # the image reads pOld after the step: the injection below does not touch it
            lib.TimeStepAdjointBornAccumulation_V(prop.p, gimg.ptr, *dps)
        else:
            lib.TimeStep(prop.p)
        if m <= nt:
            lib.InjectRegistered(prop.p, m)
        if m >= 2 and not fused:
            lib.AdjointBornAccumulation_V(prop.p, gimg.ptr, *dps)
    prop._check()
    # 4. into the shared image
    ginsu.accumulate(image, window, gimg.ptr, window.interior(nsponge, freesurface), stream)
    prop3d.synchronize(prop)
    t4 = time.perf_counter()
    out = {"born_data": born_data, "window": window, "t_setup": t1 - t0, "t_forward": t2 - t1, "t_born": t3 - t2,
           "t_adjoint": t4 - t3, "t_total": t4 - t0}
    dv.free()
    gimg.free()
    if own_snap:
        snapshots.free()
    prop3d.free(prop)
    return out


def model_shot(physics, window, spacing, dt, shot, model, nsponge=20, freesurface=True, freqQ=5.0, qMin=0.1,
               qInterior=100.0):
    """Nonlinear modelling of one shot on `window`: the receivers' P traces (nr, nt)."""
    nz, ny, nx = window.size
    dz, dy, dx = spacing
    nt = shot.wavelet.shape[1]
    nf = 1 if physics == "iso" else 2
    prop = _CLS[physics](nz=nz, ny=ny, nx=nx, nsponge=nsponge, dz=dz, dy=dy, dx=dx, dt=dt, freesurface=freesurface,
                         freqQ=freqQ, qMin=qMin, qInterior=qInterior)
    prop3d.coherence(prop, "device")
    stream = prop._lib.GetStream(prop.p)
    for name in MODEL_FIELDS[physics]:
        ginsu.extract(model[name], window, prop3d.device_ptr(prop, name), stream)
    fsi = 1 if freesurface else 0
    sz, sy, sx = _shift([shot.src[0]], [shot.src[1]], [shot.src[2]], window, spacing)
    rz, ry, rx = _shift(shot.rz, shot.ry, shot.rx, window, spacing)
    src_pts = spacetime.hickscoeffs(F32(dz), F32(dy), F32(dx), 0.0, 0.0, 0.0, nz, ny, nx, sz, sy, sx, fs_index=fsi)
    rcv_pts = spacetime.linearcoeffs(F32(dz), F32(dy), F32(dx), 0.0, 0.0, 0.0, nz, ny, nx, rz, ry, rx,
                                     fs_index=fsi)
    prop3d.set_sources(prop, [src_pts], shot.wavelet, fields=1 if nf == 1 else 3)
    prop3d.set_receivers(prop, [rcv_pts], nt)
    for it in range(1, nt + 1):
        prop._lib.TimeStepInjectExtract(prop.p, it)
    prop._check()
    out = prop3d.receiver_data(prop, shot.nr, nt)
    prop3d.free(prop)
    return out
