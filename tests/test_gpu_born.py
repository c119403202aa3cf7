"""GPU parity of the Born source injections, imaging conditions and ScaleSpatialDerivatives
against the CPU oracle (bit-exact: same operations in the same order), for caller arrays in
host memory (staged) and in device memory (torch tensors, used in place)."""
import numpy as np
import pytest

from tests.helpers import host_view, make_fields, make_model, rel_l2, smooth, to_oracle

pytestmark = pytest.mark.gpu

SHAPE = (29, 27, 33)


def _setup(oracle, kind, fs=True, seed=5):
    import wavefd_amd as w
    nx, ny, nz = SHAPE
    d = (10.0, 12.5, 8.0)
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[kind]
    ocls = {"iso": oracle.OracleIso, "vti": oracle.OracleVTI, "tti": oracle.OracleTTI}[kind]
    prop = cls(nz=nz, ny=ny, nx=nx, dz=d[2], dy=d[1], dx=d[0], dt=1e-3, nsponge=6, freesurface=fs, freqQ=10.0,
               qMin=0.5, qInterior=80.0)
    orc = ocls(nx, ny, nz, *d, dt=1e-3, fs=fs, nsponge=6, freqQ=10.0, qMin=0.5, qInterior=80.0, nthread=8)
    model = make_model(kind, SHAPE, seed)
    fields = make_fields(kind, SHAPE, seed)
    for name, arr in list(model.items()) + list(fields.items()):
        host_view(prop, name)[:] = arr.ravel()
        getattr(orc, name)[...] = arr
    return w, prop, orc


def _rand(seed, n=1):
    rng = np.random.default_rng(seed)
    out = [smooth(rng, SHAPE, 1.0, 1) for _ in range(n)]
    return out if n > 1 else out[0]


def _as(kind_of_array, arr):
    """the same values as a host (Julia-layout) numpy array or a device tensor"""
    if kind_of_array == "host":
        return np.asfortranarray(arr.transpose(2, 1, 0))  # (nz,ny,nx) F-order == same memory
    from tests.devmem import DeviceArray
    return DeviceArray(arr)


def _back(kind_of_array, a):
    if kind_of_array == "host":
        return np.ascontiguousarray(a.transpose(2, 1, 0))
    return a.numpy()


@pytest.mark.parametrize("where", ["host", "device"])
def test_iso_born_injection(oracle, where):
    w, prop, orc = _setup(oracle, "iso")
    dV, dB, P, DP = _rand(1, 4)
    w.forwardBornInjection(prop, w.Prop3DAcoIsoDenQ_DEO2_FDTD_Model_V(), {"v": _as(where, dV)},
                           {"pspace": _as(where, DP)})
    orc.forward_born_injection_V(dV, DP)
    assert np.array_equal(to_oracle(prop, "pCur", SHAPE), orc.pCur)
    w.forwardBornInjection(prop, w.Prop3DAcoIsoDenQ_DEO2_FDTD_Model_VB(), {"v": _as(where, dV), "b": _as(where, dB)},
                           {"pold": _as(where, P), "pspace": _as(where, DP)})
    orc.forward_born_injection_VB(dV, dB, P, DP)
    assert np.array_equal(to_oracle(prop, "pCur", SHAPE), orc.pCur)
    w.forwardBornInjection(prop, w.Prop3DAcoIsoDenQ_DEO2_FDTD_Model_B(), {"b": _as(where, dB)},
                           {"pold": _as(where, P), "pspace": _as(where, DP)})
    orc.forward_born_injection_B(dB, P, DP)
    assert np.array_equal(to_oracle(prop, "pCur", SHAPE), orc.pCur)


@pytest.mark.parametrize("where", ["host", "device"])
def test_iso_imaging_conditions(oracle, where):
    w, prop, orc = _setup(oracle, "iso")
    gV0, gB0, P, DP = _rand(2, 4)
    gV, gB = _as(where, gV0), _as(where, gB0)
    ogV, ogB = gV0.copy(), gB0.copy()
    ic = w.ImagingConditionStandard()
    w.adjointBornAccumulation(prop, w.Prop3DAcoIsoDenQ_DEO2_FDTD_Model_V(), ic, {"v": gV}, {"pspace": _as(where, DP)})
    orc.adjoint_born_accumulation_V(ogV, DP)
    assert np.array_equal(_back(where, gV), ogV)
    w.adjointBornAccumulation(prop, w.Prop3DAcoIsoDenQ_DEO2_FDTD_Model_VB(), ic, {"v": gV, "b": gB},
                              {"pold": _as(where, P), "pspace": _as(where, DP)})
    orc.adjoint_born_accumulation_VB(ogV, ogB, P, DP)
    assert np.array_equal(_back(where, gV), ogV)
    assert np.array_equal(_back(where, gB), ogB)
    w.adjointBornAccumulation(prop, w.Prop3DAcoIsoDenQ_DEO2_FDTD_Model_B(), ic, {"b": gB},
                              {"pold": _as(where, P), "pspace": _as(where, DP)})
    orc.adjoint_born_accumulation_B(ogB, P, DP)
    assert np.array_equal(_back(where, gB), ogB)


@pytest.mark.parametrize("kind", ["vti", "tti"])
@pytest.mark.parametrize("where", ["host", "device"])
def test_aniso_born_injection(oracle, kind, where):
    w, prop, orc = _setup(oracle, kind)
    dV, dE, dA, P, M, DP, DM = _rand(3, 7)
    mv = getattr(w, "Prop3DAco%sDenQ_DEO2_FDTD_Model_V" % kind.upper())()
    mvea = getattr(w, "Prop3DAco%sDenQ_DEO2_FDTD_Model_VEA" % kind.upper())()
    w.forwardBornInjection(prop, mv, {"v": _as(where, dV)}, {"pspace": _as(where, DP), "mspace": _as(where, DM)})
    orc.forward_born_injection_V(dV, DP, DM)
    assert np.array_equal(to_oracle(prop, "pCur", SHAPE), orc.pCur)
    assert np.array_equal(to_oracle(prop, "mCur", SHAPE), orc.mCur)
    w.forwardBornInjection(prop, mvea, {"v": _as(where, dV), "ϵ": _as(where, dE), "η": _as(where, dA)},
                           {"pold": _as(where, P), "mold": _as(where, M), "pspace": _as(where, DP),
                            "mspace": _as(where, DM)})
    orc.forward_born_injection_VEA(dV, dE, dA, P, M, DP, DM)
    assert np.array_equal(to_oracle(prop, "pCur", SHAPE), orc.pCur)
    assert np.array_equal(to_oracle(prop, "mCur", SHAPE), orc.mCur)


@pytest.mark.parametrize("kind", ["vti", "tti"])
@pytest.mark.parametrize("where", ["host", "device"])
def test_aniso_imaging_conditions(oracle, kind, where):
    w, prop, orc = _setup(oracle, kind)
    gV0, gE0, gA0, P, M, DP, DM = _rand(4, 7)
    gV, gE, gA = _as(where, gV0), _as(where, gE0), _as(where, gA0)
    ogV, ogE, ogA = gV0.copy(), gE0.copy(), gA0.copy()
    ic = w.ImagingConditionStandard()
    mv = getattr(w, "Prop3DAco%sDenQ_DEO2_FDTD_Model_V" % kind.upper())()
    mvea = getattr(w, "Prop3DAco%sDenQ_DEO2_FDTD_Model_VEA" % kind.upper())()
    w.adjointBornAccumulation(prop, mv, ic, {"v": gV}, {"pspace": _as(where, DP), "mspace": _as(where, DM)})
    orc.adjoint_born_accumulation_V(ogV, DP, DM)
    assert np.array_equal(_back(where, gV), ogV)
    w.adjointBornAccumulation(prop, mvea, ic, {"v": gV, "ϵ": gE, "η": gA},
                              {"pold": _as(where, P), "mold": _as(where, M), "pspace": _as(where, DP),
                               "mspace": _as(where, DM)})
    orc.adjoint_born_accumulation_VEA(ogV, ogE, ogA, P, M, DP, DM)
    assert np.array_equal(_back(where, gV), ogV)
    assert np.array_equal(_back(where, gE), ogE)
    assert np.array_equal(_back(where, gA), ogA)


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_scale_spatial_derivatives(oracle, kind):
    w, prop, orc = _setup(oracle, kind)
    for _ in range(2):
        w.propagateforward(prop)
        orc.time_step()
    w.scale_spatial_derivatives(prop)
    orc.scale_spatial_derivatives()
    assert np.array_equal(to_oracle(prop, "pSpace", SHAPE), orc.pSpace)
    if kind != "iso":
        assert np.array_equal(to_oracle(prop, "mSpace", SHAPE), orc.mSpace)


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("isfwi", [True, False])
def test_wavefield_separation_matches_fft_oracle(oracle, kind, isfwi):
    """hipFFT (float32) vs numpy pocketfft (float64) restatement; FFTW bits are not available."""
    w, prop, orc = _setup(oracle, kind)
    g0, DP, DM = _rand(6, 3)
    g = _as("host", g0)
    og = g0.copy()
    ic = w.ImagingConditionWaveFieldSeparationFWI() if isfwi else w.ImagingConditionWaveFieldSeparationRTM()
    mv = getattr(w, "Prop3DAco%sDenQ_DEO2_FDTD_Model_V" % ("Iso" if kind == "iso" else kind.upper()))()
    wf = {"pspace": _as("host", DP)}
    if kind != "iso":
        wf["mspace"] = _as("host", DM)
    w.adjointBornAccumulation(prop, mv, ic, {"v": g}, wf)
    if kind == "iso":
        orc.adjoint_born_accumulation_wavefieldsep(og, DP, isfwi)
    else:
        orc.adjoint_born_accumulation_wavefieldsep_V(og, DP, DM, isfwi)
    got = _back("host", g)
    assert rel_l2(got - g0, og - g0) < 2e-5


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_imaging_condition_heuristics(kind):
    """test/imgcondition.jl:77-130 restated: a horizontal bar correlated with itself."""
    import wavefd_amd as w
    nz, ny, nx = 50, 30, 40
    z1, z2 = 20, 30
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[kind]
    prop = cls(freesurface=False, nz=nz, ny=ny, nx=nx, dz=10.0, dy=10.0, dx=10.0, dt=1e-3)
    w.B(prop)[...] = 1
    w.V(prop)[...] = 1
    w.POld(prop)[z1:z2, :, :] = 1
    if kind != "iso":
        w.MOld(prop)[z1:z2, :, :] = 2
    wp = np.zeros((nz, ny, nx), np.float32, order="F")
    wm = np.zeros((nz, ny, nx), np.float32, order="F")
    wp[z1:z2] = 1
    wm[z1:z2] = 2
    gstd, gfwi, grtm = (np.zeros((nz, ny, nx), np.float32, order="F") for _ in range(3))
    mv = getattr(w, "Prop3DAco%sDenQ_DEO2_FDTD_Model_V" % ("Iso" if kind == "iso" else kind.upper()))()
    wf = {"pspace": wp, "mspace": wm}
    w.adjointBornAccumulation(prop, mv, w.ImagingConditionStandard(), {"v": gstd}, wf)
    w.adjointBornAccumulation(prop, mv, w.ImagingConditionWaveFieldSeparationFWI(), {"v": gfwi}, wf)
    w.adjointBornAccumulation(prop, mv, w.ImagingConditionWaveFieldSeparationRTM(), {"v": grtm}, wf)
    assert gstd.min() >= 0
    assert gstd.max() >= 0
    assert gfwi.min() > 0
    assert gfwi.max() > 0
    assert grtm.min() < 0
    assert grtm.max() > 0
    assert np.sign(gfwi).sum() > np.sign(gstd).sum()
    assert float(np.dot(gstd.ravel(), gfwi.ravel())) > float(np.dot(gstd.ravel(), grtm.ravel()))
