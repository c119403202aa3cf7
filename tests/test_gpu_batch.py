"""GPU: several grids stepped by one launch (TimeStepBatch) equal the same grids stepped one by
one, bit for bit, for every epilogue (plain step, save DP, Born injection, imaging, and Iso's
coupled background + Born step), for Iso, VTI and TTI, every tile shape (the wide tiles through
the wide batch kernel), grids of different sizes with and without a free surface; the batch's
argument checks; and the batched Born driver (born_shots) gives the same Born data and image as
the shots run one after another (born_shot)."""
import ctypes

import numpy as np
import pytest

from tests.helpers import host_view, make_fields, make_model, smooth

pytestmark = pytest.mark.gpu

F32 = np.float32
CLS = {"iso": "Prop3DAcoIsoDenQ_DEO2_FDTD", "vti": "Prop3DAcoVTIDenQ_DEO2_FDTD", "tti": "Prop3DAcoTTIDenQ_DEO2_FDTD"}
GRIDS = [((37, 41, 45), True), ((20, 23, 61), False), ((30, 29, 33), True)]


def _prop(kind, shape, fs, seed, tile):
    import wavefd_amd as w
    nx, ny, nz = shape
    p = getattr(w, CLS[kind])(nz=nz, ny=ny, nx=nx, dz=10.0, dy=11.0, dx=9.0, dt=1e-3, nsponge=6, freesurface=fs,
                              freqQ=10.0, qMin=0.5, qInterior=80.0)
    for name, arr in make_model(kind, shape, seed).items():
        host_view(p, name)[:] = arr.ravel()
    for name, arr in make_fields(kind, shape, seed + 1).items():
        host_view(p, name)[:] = arr.ravel()
    w.coherence(p, "device")
    w.sync_to_device(p)
    w.set_fused_tile(p, tile)
    return p


def _dev(arr):
    from wavefd_amd import ginsu
    b = ginsu.DeviceBuffer(arr.nbytes)
    b.upload(np.ascontiguousarray(arr, F32))
    return b


def _fields(kind):
    return ("pCur", "pOld") + (("mCur", "mOld") if kind != "iso" else ())


def _dev_field(p, name, n):
    import wavefd_amd as w
    from wavefd_amd import ginsu
    out = np.zeros(n, F32)
    w.synchronize(p)
    ginsu.copy(out.ctypes.data, w.device_ptr(p, name), 4 * n)
    ginsu.sync()
    return out


# epilogue 4, the coupled Born step, is Iso's (VTI rejects it: test_batch_argument_checks)
@pytest.mark.parametrize("kind,epi", [(k, e) for k in ("iso", "vti", "tti") for e in range(5) if e < 4 or k == "iso"])
@pytest.mark.parametrize("tile", ["64x16", "32x32", "wide", "wide2", "auto"])
def test_batch_equals_one_by_one(kind, tile, epi):
    import wavefd_amd as w
    from wavefd_amd import _lib
    nf = 1 if kind == "iso" else 2
    a = [_prop(kind, s, fs, 3 + i, tile) for i, (s, fs) in enumerate(GRIDS)]
    b = [_prop(kind, s, fs, 3 + i, tile) for i, (s, fs) in enumerate(GRIDS)]
    rng = np.random.default_rng(epi)
    # per grid: the epilogue's arrays (a0, a1[, a2]), one device copy for each side
    arrs_a, arrs_b, keep = [], [], []
    for s, _ in GRIDS:
        n = int(np.prod(s))
        if epi == 0:
            host = []
        elif epi == 1:
            host = [np.zeros(n, F32) for _ in range(nf)]
        elif epi == 4:  # dV, DP out
            host = [smooth(rng, s, 40.0, 2).ravel(), np.zeros(n, F32)]
        else:
            host = [smooth(rng, s, 40.0 if epi == 2 else 1.0, 2).ravel()] + [smooth(rng, s, 1.0, 2).ravel()
                                                                            for _ in range(nf)]
        da, db = [_dev(h) for h in host], [_dev(h) for h in host]
        keep += da + db
        arrs_a.append(da)
        arrs_b.append(db)
    lib = a[0]._lib
    # one by one through the single-handle entry points
    for p, da in zip(a, arrs_a):
        ptrs = [d.ptr for d in da]
        if epi == 0:
            lib.TimeStep(p.p)
        elif epi == 1:
            assert lib.TimeStepSaveDP(p.p, *ptrs) == 0
        elif epi == 2:
            assert lib.TimeStepBornInjection_V(p.p, *ptrs) == 0
        elif epi == 3:
            assert lib.TimeStepAdjointBornAccumulation_V(p.p, *ptrs) == 0
        else:
            assert lib.TimeStepBornCoupled_V(p.p, *ptrs) == 0
    # all at once
    n = len(b)
    hs = (ctypes.c_void_p * n)(*[p.p for p in b])
    cols = [None, None, None]
    for k in range(len(arrs_b[0])):
        cols[k] = (ctypes.c_void_p * n)(*[arrs_b[i][k].ptr for i in range(n)])
    assert lib.TimeStepBatch(hs, n, epi, *cols) == 0, _lib.last_error(kind)
    for p in a + b:
        w.sync_to_host(p)
    for i, (s, _) in enumerate(GRIDS):
        for nm in _fields(kind):
            assert np.array_equal(host_view(b[i], nm), host_view(a[i], nm)), (i, nm)
        if epi == 4:  # the Born field
            for nm in ("mCur", "mOld"):
                n = int(np.prod(s))
                assert np.array_equal(_dev_field(b[i], nm, n), _dev_field(a[i], nm, n)), (i, nm)
        if epi == 0:
            assert np.array_equal(host_view(b[i], "pSpace"), host_view(a[i], "pSpace"))
        for k, (da, db) in enumerate(zip(arrs_a[i], arrs_b[i])):
            ga = da.download(np.zeros(int(np.prod(s)), F32))
            gb = db.download(np.zeros(int(np.prod(s)), F32))
            assert np.array_equal(ga, gb), (i, k)
    for p in a + b:
        w.free(p)
    for d in keep:
        d.free()


def test_batch_argument_checks():
    import wavefd_amd as w
    from wavefd_amd import _lib
    a = [_prop("iso", s, fs, 3, "auto") for s, fs in GRIDS[:2]]
    lib = a[0]._lib

    def call(props, n, epi=0, cols=(None, None, None)):
        hs = (ctypes.c_void_p * max(1, len(props)))(*[p.p for p in props])
        return lib.TimeStepBatch(hs, n, epi, *cols)

    assert call(a, 0) == -1 and "between 1 and 8" in _lib.last_error("iso")
    assert call([a[0], a[0]], 2) == -1 and "twice" in _lib.last_error("iso")
    assert call(a, 2, 5) == -1 and "epilogue" in _lib.last_error("iso")
    v = [_prop("vti", s, fs, 3, "auto") for s, fs in GRIDS[:2]]
    hv = (ctypes.c_void_p * 2)(*[p.p for p in v])
    assert v[0]._lib.TimeStepBatch(hv, 2, 4, None, None, None) == -1 and "epilogue" in _lib.last_error("vti")
    v[0]._lib.ClearError()
    for p in v:
        w.free(p)
    host = np.zeros(int(np.prod(GRIDS[0][0])), F32)
    col = (ctypes.c_void_p * 2)(host.ctypes.data, host.ctypes.data)
    assert call(a, 2, 1, (col, None, None)) == -1 and "device memory" in _lib.last_error("iso")
    assert call(a, 2, 1) == -1 and "missing" in _lib.last_error("iso")
    w.coherence(a[1], "coherent")
    assert call(a, 2) == -1 and "DEVICE" in _lib.last_error("iso")
    lib.ClearError()
    for p in a:
        w.free(p)


@pytest.mark.parametrize("kind", ["iso", "vti"])
def test_born_shots_batched_equal_sequential(kind):
    from wavefd_amd import born, ginsu, spacetime
    shape = (56, 50, 60)  # big (nx, ny, nz)
    nx, ny, nz = shape
    m = make_model(kind, shape, 21)
    dev = {k: _dev(m[k]) for k in born.MODEL_FIELDS[kind]}
    rng = np.random.default_rng(22)
    dv = smooth(rng, shape, 40.0, 2).astype(F32)
    dvb = _dev(dv)
    nt = 80
    wav = spacetime.causal_ricker(25.0, 1e-3 * np.arange(nt))
    big = (nz, ny, nx)
    windows, shots = [], []
    for k, (cy, cx) in enumerate([(200.0, 220.0), (300.0, 330.0), (250.0, 280.0)]):
        sh = born.BornShot((45.0, cy, cx), np.full(5, 35.0), np.linspace(cy - 60, cy + 60, 5),
                           np.linspace(cx - 50, cx + 70, 5), wav)
        windows.append(ginsu.window_for_shot(big, (10.0, 10.0, 10.0), (0.0, 0.0, 0.0), sh.src, sh.rz, sh.ry, sh.rx,
                                             aperture=(40.0 + 20 * k, 30.0), pad=8))
        shots.append(sh)
    assert len({w.size for w in windows}) > 1  # uneven windows
    model = {k: b.ptr for k, b in dev.items()}
    img_seq = ginsu.DeviceBuffer(dv.nbytes)
    img_seq.zero()
    seq = [born.born_shot(kind, w, (10.0, 10.0, 10.0), 1e-3, s, model, dvb.ptr, img_seq.ptr, nsponge=8)["born_data"]
           for w, s in zip(windows, shots)]
    img_bat = ginsu.DeviceBuffer(dv.nbytes)
    img_bat.zero()
    bat = born.born_shots(kind, windows, (10.0, 10.0, 10.0), 1e-3, shots, model, dvb.ptr, img_bat.ptr, nsponge=8,
                          batch=3)
    for x, y in zip(seq, bat):
        assert np.array_equal(x, y)
        assert np.abs(x).max() > 0
    gs = img_seq.download(np.zeros(shape, F32))
    gb = img_bat.download(np.zeros(shape, F32))
    assert np.array_equal(gs, gb)
    assert np.abs(gs).max() > 0
