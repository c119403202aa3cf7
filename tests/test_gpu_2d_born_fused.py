"""GPU: the 2D libraries' fused Born steps equal the reference's two-call pairs bit for bit, on
every cell, three steps in a row:

TimeStepSaveDP          = TimeStep; ScaleSpatialDerivatives      (DP, DM vs the scaled pSpace, mSpace)
TimeStepBornInjection_V = TimeStep; ForwardBornInjection_V       (new pCur, mCur)
TimeStepAdjointBornAccumulation_V = TimeStep; AdjointBornAccumulation_V (g, new pCur, mCur)

for Iso, VTI and TTI, with and without a free surface, with the fused kernel and with the
two-pass step (SetFusedStep 0), on a grid whose second z-strip is ragged (300 cells: 249 + 51)."""
import numpy as np
import pytest

from tests.helpers import smooth

pytestmark = pytest.mark.gpu

F32 = np.float32
SHAPE = (37, 300)  # (nx, nz)
CLS = {"iso": "Prop2DAcoIsoDenQ_DEO2_FDTD", "vti": "Prop2DAcoVTIDenQ_DEO2_FDTD", "tti": "Prop2DAcoTTIDenQ_DEO2_FDTD"}


def _prop(kind, fs, fused):
    import wavefd_amd.prop2d as w2
    nx, nz = SHAPE
    p = getattr(w2, CLS[kind])(nz=nz, nx=nx, dz=8.0, dx=10.0, dt=1e-3, nsponge=6, freesurface=fs, freqQ=10.0,
                               qMin=0.5, qInterior=80.0)
    rng = np.random.default_rng(3)
    arrs = {"V": 2000 + smooth(rng, SHAPE, 400.0, 2), "B": 0.8 + smooth(rng, SHAPE, 0.25, 2),
            "PCur": smooth(rng, SHAPE, 1.0, 2), "POld": smooth(rng, SHAPE, 1.0, 2)}
    if kind != "iso":
        arrs.update(Eps=0.1 + smooth(rng, SHAPE, 0.08, 2), Eta=0.05 + smooth(rng, SHAPE, 0.04, 2),
                    F=np.full(SHAPE, 0.85), MCur=smooth(rng, SHAPE, 1.0, 2), MOld=smooth(rng, SHAPE, 1.0, 2))
    if kind == "tti":
        th = smooth(rng, SHAPE, 0.7, 2)
        arrs.update(SinTheta=np.sin(th), CosTheta=np.cos(th))
    for name, a in arrs.items():
        getattr(w2, name)(p)[...] = np.asarray(a, F32).T  # (nx, nz) C order == Julia (nz, nx)
    w2.set_fused_step(p, fused)
    return p


def _flat(p, name):
    import wavefd_amd.prop2d as w2
    return np.asarray(getattr(w2, name)(p)).ravel(order="F").copy()


def _fields(kind):
    return ("PCur", "POld") + (("MCur", "MOld") if kind != "iso" else ())


def _arrs(kind, seed, amp=1.0):
    rng = np.random.default_rng(seed)
    return [smooth(rng, SHAPE, amp, 2).ravel().copy() for _ in range(1 if kind == "iso" else 2)]


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("fused", [True, False])
def test_save_dp(kind, fs, fused):
    a, b = _prop(kind, fs, fused), _prop(kind, fs, fused)
    for _ in range(3):
        a._lib.TimeStep(a.p)
        a._lib.ScaleSpatialDerivatives(a.p)
        out = [np.zeros(np.prod(SHAPE), F32) for _ in range(1 if kind == "iso" else 2)]
        assert b._lib.TimeStepSaveDP(b.p, *[o.ctypes.data for o in out]) == 0
        assert np.array_equal(out[0], _flat(a, "PSpace"))
        if kind != "iso":
            assert np.array_equal(out[1], _flat(a, "MSpace"))
        for nm in _fields(kind):
            assert np.array_equal(_flat(b, nm), _flat(a, nm)), nm
    assert np.abs(out[0]).max() > 0


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("fused", [True, False])
def test_born_injection(kind, fs, fused):
    a, b = _prop(kind, fs, fused), _prop(kind, fs, fused)
    for step in range(3):
        dv = _arrs("iso", 10 + step, 40.0)[0]
        dps = _arrs(kind, 20 + step)
        ptrs = [dv.ctypes.data] + [d.ctypes.data for d in dps]
        a._lib.TimeStep(a.p)
        a._lib.ForwardBornInjection_V(a.p, *ptrs)
        assert b._lib.TimeStepBornInjection_V(b.p, *ptrs) == 0
        for nm in _fields(kind):
            assert np.array_equal(_flat(b, nm), _flat(a, nm)), (nm, step)


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("fused", [True, False])
def test_imaging(kind, fs, fused):
    a, b = _prop(kind, fs, fused), _prop(kind, fs, fused)
    g0 = _arrs("iso", 5)[0]
    ga, gb = g0.copy(), g0.copy()
    for step in range(3):
        dps = _arrs(kind, 30 + step)
        a._lib.TimeStep(a.p)
        a._lib.AdjointBornAccumulation_V(a.p, ga.ctypes.data, *[d.ctypes.data for d in dps])
        assert b._lib.TimeStepAdjointBornAccumulation_V(b.p, gb.ctypes.data, *[d.ctypes.data for d in dps]) == 0
        assert np.array_equal(gb, ga), step
        for nm in _fields(kind):
            assert np.array_equal(_flat(b, nm), _flat(a, nm)), (nm, step)
    assert not np.array_equal(ga, g0)
