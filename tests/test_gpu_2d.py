"""GPU: the 2D Iso library (libprop2DAcoIsoDenQ_DEO2_FDTD) against the 2D oracle
(oracle/oracle2d.py): the Q profile, multi-step wavefields and the Born sources and imaging
conditions bit for bit, with and without a free surface, on grids whose interiors are not
multiples of the kernels' blocks; wavefield-separation imaging against numpy's FFT (float64)
to 2e-5 relative (FFTW's float32 bits are not available)."""
import numpy as np
import pytest

from oracle import oracle2d as o2
from tests.helpers import smooth

pytestmark = pytest.mark.gpu

F32 = np.float32
SHAPES = [(37, 45), (21, 70)]  # (nx, nz)


def _pair(shape, fs, seed):
    import wavefd_amd.prop2d as w2
    nx, nz = shape
    prop = w2.Prop2DAcoIsoDenQ_DEO2_FDTD(nz=nz, nx=nx, dz=8.0, dx=10.0, dt=1e-3, nsponge=6, freesurface=fs,
                                         freqQ=10.0, qMin=0.5, qInterior=80.0)
    orc = o2.Oracle2DIso(nx, nz, 10.0, 8.0, 1e-3, fs=fs, nsponge=6, freqQ=10.0, qMin=0.5, qInterior=80.0)
    rng = np.random.default_rng(seed)
    arrs = {"v": (2000 + smooth(rng, shape, 400.0, 2)).astype(F32), "b": (0.8 + smooth(rng, shape, 0.25, 2)).astype(F32),
            "pCur": smooth(rng, shape, 1.0, 2), "pOld": smooth(rng, shape, 1.0, 2)}
    for name, a in arrs.items():
        getattr(w2, {"v": "V", "b": "B", "pCur": "PCur", "pOld": "POld"}[name])(prop)[...] = a.T
        getattr(orc, name)[...] = a
    return prop, orc


def _get(prop, name):
    import wavefd_amd.prop2d as w2
    return np.ascontiguousarray(getattr(w2, name)(prop).T)  # (nz, nx) Fortran -> (nx, nz) C


@pytest.mark.parametrize("fs", [False, True])
def test_q_profile(fs):
    prop, orc = _pair((33, 41), fs, 1)
    assert np.array_equal(_get(prop, "DtOmegaInvQ"), orc.q)


@pytest.mark.parametrize("fs", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_steps_bit_exact(fs, shape):
    import wavefd_amd.prop2d as w2
    prop, orc = _pair(shape, fs, 7)
    for _ in range(6):
        w2.propagateforward(prop)
        orc.time_step()
    for gname, oname in (("PCur", "pCur"), ("POld", "pOld"), ("PSpace", "pSpace")):
        assert np.array_equal(_get(prop, gname), getattr(orc, oname)), gname


@pytest.mark.parametrize("fs", [False, True])
def test_many_steps(fs):
    import wavefd_amd.prop2d as w2
    prop, orc = _pair((48, 52), fs, 9)
    for _ in range(300):
        w2.propagateforward(prop)
        orc.time_step()
    assert np.array_equal(_get(prop, "PCur"), orc.pCur)


@pytest.mark.parametrize("which", ["V", "VB", "B"])
@pytest.mark.parametrize("fs", [False, True])
def test_born_source_bit_exact(which, fs):
    import wavefd_amd.prop2d as w2
    shape = (37, 45)
    prop, orc = _pair(shape, fs, 11)
    rng = np.random.default_rng(12)
    dv, db = smooth(rng, shape, 50.0, 2), smooth(rng, shape, 0.02, 2)
    P, DP = smooth(rng, shape, 1.0, 2), smooth(rng, shape, 1.0, 2)
    cls = {"V": w2.Prop2DAcoIsoDenQ_DEO2_FDTD_Model_V, "VB": w2.Prop2DAcoIsoDenQ_DEO2_FDTD_Model_VB,
           "B": w2.Prop2DAcoIsoDenQ_DEO2_FDTD_Model_B}[which]
    # the library takes (nz, nx) column-major arrays: the (nx, nz) C-order arrays' memory
    w2.forwardBornInjection(prop, cls(), {"v": dv, "b": db}, {"pold": P, "pspace": DP})
    if which == "V":
        orc.forward_born_injection_V(dv, DP)
    elif which == "VB":
        orc.forward_born_injection_VB(dv, db, P, DP)
    else:
        orc.forward_born_injection_B(db, P, DP)
    assert np.array_equal(_get(prop, "PCur"), orc.pCur)


@pytest.mark.parametrize("which", ["V", "VB", "B"])
@pytest.mark.parametrize("fs", [False, True])
def test_imaging_bit_exact(which, fs):
    import wavefd_amd.prop2d as w2
    shape = (37, 45)
    prop, orc = _pair(shape, fs, 13)
    rng = np.random.default_rng(14)
    P, DP = smooth(rng, shape, 1.0, 2), smooth(rng, shape, 1.0, 2)
    gv, gb = smooth(rng, shape, 1.0, 2), smooth(rng, shape, 1.0, 2)
    gv2, gb2 = gv.copy(), gb.copy()
    cls = {"V": w2.Prop2DAcoIsoDenQ_DEO2_FDTD_Model_V, "VB": w2.Prop2DAcoIsoDenQ_DEO2_FDTD_Model_VB,
           "B": w2.Prop2DAcoIsoDenQ_DEO2_FDTD_Model_B}[which]
    w2.adjointBornAccumulation(prop, cls(), w2.ImagingConditionStandard(), {"v": gv, "b": gb},
                               {"pold": P, "pspace": DP})
    if which == "V":
        orc.adjoint_born_accumulation_V(gv2, DP)
    elif which == "VB":
        orc.adjoint_born_accumulation_VB(gv2, gb2, P, DP)
    else:
        orc.adjoint_born_accumulation_B(gb2, P, DP)
    assert np.array_equal(gv, gv2)
    assert np.array_equal(gb, gb2)


@pytest.mark.parametrize("fwi", [True, False])
def test_wavefield_separation(fwi):
    import wavefd_amd.prop2d as w2
    shape = (23, 40)
    prop, orc = _pair(shape, True, 15)
    rng = np.random.default_rng(16)
    DP = smooth(rng, shape, 1.0, 2)
    g = np.zeros(shape, F32)
    ic = w2.ImagingConditionWaveFieldSeparationFWI() if fwi else w2.ImagingConditionWaveFieldSeparationRTM()
    w2.adjointBornAccumulation(prop, w2.Prop2DAcoIsoDenQ_DEO2_FDTD_Model_V(), ic, {"v": g}, {"pspace": DP})
    # numpy restatement in float64 (src/prop2DAcoIsoDenQ_DEO2_FDTD.h:584-672): per x column,
    # zero-pad to 2 nz, forward FFT with sign +1, keep the down-going (nonlinear) half, the
    # adjoint's down-going (FWI) or up-going (RTM) half, inverse, real part of the product
    nx, nz = shape
    nfft = 2 * nz
    v, b, pold = orc.v.astype(np.float64), orc.b.astype(np.float64), orc.pOld.astype(np.float64)
    want = np.zeros(shape)
    kadj = 0 if fwi else nfft // 2
    for kx in range(nx):
        a = np.zeros(nfft, complex)
        r = np.zeros(nfft, complex)
        a[:nz] = DP[kx] / nfft
        r[:nz] = pold[kx] / nfft
        A = np.fft.ifft(a) * nfft  # sign +1
        R = np.fft.ifft(r) * nfft
        A[nfft // 2 + 1:] = 0
        R[kadj + 1:kadj + nfft // 2] = 0
        a2 = np.fft.fft(A)  # sign -1
        r2 = np.fft.fft(R)
        want[kx] = 2 * b[kx] / v[kx] ** 3 * np.real(a2[:nz] * r2[:nz])
    err = np.linalg.norm(g - want) / np.linalg.norm(want)
    assert err <= 2e-5, err
