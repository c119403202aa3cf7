"""GPU: the fused Born steps equal the reference's two-call pairs bit for bit, on every cell.

TimeStepSaveDP          = TimeStep; ScaleSpatialDerivatives      (DP, DM vs the scaled pSpace, mSpace)
TimeStepBornInjection_V = TimeStep; ForwardBornInjection_V       (new pCur, mCur)
TimeStepAdjointBornAccumulation_V = TimeStep; AdjointBornAccumulation_V (g, new pCur, mCur)

for Iso, VTI and TTI, with and without a free surface, on grids whose tiles are ragged (and whose last tiles write the annulus from their halo), with
the fused kernel's tile shapes (64x16, 32x32, wide, wide2) and with the two-pass fallback
(SetFusedStep 0); and the Born driver gives the same Born data and image with the fused pairs
as with the separate calls.
"""
import numpy as np
import pytest

from tests.helpers import host_view, make_fields, make_model, smooth

pytestmark = pytest.mark.gpu

F32 = np.float32
CLS = {"iso": "Prop3DAcoIsoDenQ_DEO2_FDTD", "vti": "Prop3DAcoVTIDenQ_DEO2_FDTD", "tti": "Prop3DAcoTTIDenQ_DEO2_FDTD"}
SHAPE = (37, 41, 45)
# the last two: the last tile along z and y writes the grid's last (annulus) cells from its upper
# halo (a few cells past a multiple of the outputs per tile for every tile shape)
SHAPES = [SHAPE, (21, 37, 60), (23, 28, 53)]


@pytest.fixture(params=SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
def shape(request):
    return request.param


def _prop(kind, fs, fused, shape=SHAPE):
    import wavefd_amd as w
    nx, ny, nz = shape
    p = getattr(w, CLS[kind])(nz=nz, ny=ny, nx=nx, dz=10.0, dy=11.0, dx=9.0, dt=1e-3, nsponge=6, freesurface=fs,
                              freqQ=10.0, qMin=0.5, qInterior=80.0)
    for name, arr in make_model(kind, shape, 3).items():
        host_view(p, name)[:] = arr.ravel()
    for name, arr in make_fields(kind, shape, 5).items():
        host_view(p, name)[:] = arr.ravel()
    p._lib.SetFusedStep(p.p, 1 if fused else 0)
    if fused:
        assert p._lib.SetFusedTile(p.p, fused) == 0
    return p


def _arrs(kind, seed, shape=SHAPE):
    rng = np.random.default_rng(seed)
    n = 1 if kind == "iso" else 2
    return [smooth(rng, shape, 1.0, 2).ravel().copy() for _ in range(n)]


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("fused", [1, 2, 3, 4, 0])  # tile 64x16, 32x32, wide, wide2, two-pass
def test_save_dp(kind, fs, fused, shape):
    a, b = _prop(kind, fs, fused, shape), _prop(kind, fs, fused, shape)
    a._lib.TimeStep(a.p)
    a._lib.ScaleSpatialDerivatives(a.p)
    out = [np.zeros(np.prod(shape), F32) for _ in range(1 if kind == "iso" else 2)]
    assert b._lib.TimeStepSaveDP(b.p, *[o.ctypes.data for o in out]) == 0
    assert np.array_equal(out[0], host_view(a, "pSpace"))
    if kind != "iso":
        assert np.array_equal(out[1], host_view(a, "mSpace"))
    for nm in ("pCur", "pOld") + (("mCur", "mOld") if kind != "iso" else ()):
        assert np.array_equal(host_view(b, nm), host_view(a, nm)), nm


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("fused", [1, 2, 3, 4, 0])  # tile 64x16, 32x32, wide, wide2, two-pass
def test_born_injection(kind, fs, fused, shape):
    a, b = _prop(kind, fs, fused, shape), _prop(kind, fs, fused, shape)
    rng = np.random.default_rng(8)
    dv = smooth(rng, shape, 50.0, 2).ravel().copy()
    dps = _arrs(kind, 9, shape)
    a._lib.TimeStep(a.p)
    a._lib.ForwardBornInjection_V(a.p, dv.ctypes.data, *[d.ctypes.data for d in dps])
    assert b._lib.TimeStepBornInjection_V(b.p, dv.ctypes.data, *[d.ctypes.data for d in dps]) == 0
    for nm in ("pCur", "pOld") + (("mCur", "mOld") if kind != "iso" else ()):
        assert np.array_equal(host_view(b, nm), host_view(a, nm)), nm


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("fused", [1, 2, 3, 4, 0])  # tile 64x16, 32x32, wide, wide2, two-pass
def test_adjoint_accumulation(kind, fs, fused, shape):
    a, b = _prop(kind, fs, fused, shape), _prop(kind, fs, fused, shape)
    rng = np.random.default_rng(10)
    g0 = smooth(rng, shape, 1.0, 2).ravel().copy()
    dps = _arrs(kind, 11, shape)
    ga, gb = g0.copy(), g0.copy()
    a._lib.TimeStep(a.p)
    a._lib.AdjointBornAccumulation_V(a.p, ga.ctypes.data, *[d.ctypes.data for d in dps])
    assert b._lib.TimeStepAdjointBornAccumulation_V(b.p, gb.ctypes.data, *[d.ctypes.data for d in dps]) == 0
    assert np.array_equal(gb, ga)
    assert not np.array_equal(ga, g0)
    for nm in ("pCur", "pOld") + (("mCur", "mOld") if kind != "iso" else ()):
        assert np.array_equal(host_view(b, nm), host_view(a, nm)), nm


@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("fused", [1, 2, 3, 4, 0])  # tile 64x16, 32x32, wide, wide2, two-pass
def test_coupled_step(fs, fused, shape):
    """TimeStepBornCoupled_V on one handle = TimeStepSaveDP on a background handle followed by
    TimeStepBornInjection_V on a handle holding the Born field, step after step"""
    import wavefd_amd as w
    a, b, c = _prop("iso", fs, fused, shape), _prop("iso", fs, fused, shape), _prop("iso", fs, fused, shape)
    n = int(np.prod(shape))
    for nm in ("pCur", "pOld"):  # the Born field starts at zero, like the coupled handle's
        host_view(b, nm)[:] = 0.0
    rng = np.random.default_rng(12)
    dv = smooth(rng, shape, 50.0, 2).ravel().copy()
    for step in range(3):
        dpa, dpc = np.zeros(n, F32), np.zeros(n, F32)
        assert a._lib.TimeStepSaveDP(a.p, dpa.ctypes.data) == 0
        assert b._lib.TimeStepBornInjection_V(b.p, dv.ctypes.data, dpa.ctypes.data) == 0
        assert c._lib.TimeStepBornCoupled_V(c.p, dv.ctypes.data, dpc.ctypes.data) == 0
        assert np.array_equal(dpa, dpc), step
        for nm in ("pCur", "pOld"):
            assert np.array_equal(host_view(c, nm), host_view(a, nm)), (step, nm)
    from wavefd_amd import ginsu
    for nm, dn in (("pCur", "mCur"), ("pOld", "mOld")):  # the Born field, kept as mCur/mOld
        born = np.zeros(n, F32)
        ptr = w.device_ptr(c, dn)
        ginsu.copy(born.ctypes.data, ptr, born.nbytes)
        ginsu.sync()
        assert np.array_equal(born, host_view(b, nm)), nm
    assert np.abs(host_view(b, "pCur")).max() > 0


@pytest.mark.parametrize("kind", ["iso", "vti"])
def test_driver_fused_equals_separate(kind):
    from wavefd_amd import born, ginsu, spacetime
    shape = (44, 42, 48)
    nx, ny, nz = shape
    m = make_model(kind, shape, 17)
    bufs = {}
    for k in born.MODEL_FIELDS[kind]:
        bufs[k] = ginsu.DeviceBuffer(m[k].nbytes)
        bufs[k].upload(m[k])
    rng = np.random.default_rng(19)
    dv = smooth(rng, shape, 40.0, 2).astype(F32)
    dvb = ginsu.DeviceBuffer(dv.nbytes)
    dvb.upload(dv)
    nt = 90
    shot = born.BornShot((45.0, 200.0, 230.0), np.full(7, 35.0), np.linspace(90, 320, 7), np.linspace(100, 370, 7),
                         spacetime.causal_ricker(25.0, 1e-3 * np.arange(nt)))
    w = ginsu.Window((nz, ny, nx), (0, 0, 0), (nz, ny, nx))
    res = []
    modes = [(True, None), (False, None)] + ([(True, True)] if kind == "iso" else [])
    for fused, coupled in modes:  # fused passes (the default with exact storage), separate calls, Iso coupled
        img = ginsu.DeviceBuffer(dv.nbytes)
        img.zero()
        out = born.born_shot(kind, w, (10.0, 10.0, 10.0), 1e-3, shot, {k: q.ptr for k, q in bufs.items()}, dvb.ptr,
                             img.ptr, nsponge=8, fused=fused, coupled=coupled)
        res.append((out["born_data"], img.download(np.zeros(shape, F32))))
    for r in res[1:]:
        assert np.array_equal(res[0][0], r[0])
        assert np.array_equal(res[0][1], r[1])
    assert np.abs(res[0][1]).max() > 0
