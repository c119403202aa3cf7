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


# ---- 2D VTI / TTI ---------------------------------------------------------------------------

_GET = {"v": "V", "b": "B", "eps": "Eps", "eta": "Eta", "f": "F", "sinTheta": "SinTheta", "cosTheta": "CosTheta",
        "pCur": "PCur", "pOld": "POld", "mCur": "MCur", "mOld": "MOld", "pSpace": "PSpace", "mSpace": "MSpace",
        "q": "DtOmegaInvQ"}


def _apair(kind, shape, fs, seed):
    import wavefd_amd.prop2d as w2
    from tests.helpers import make_model
    nx, nz = shape
    cls = {"vti": w2.Prop2DAcoVTIDenQ_DEO2_FDTD, "tti": w2.Prop2DAcoTTIDenQ_DEO2_FDTD}[kind]
    prop = cls(nz=nz, nx=nx, dz=8.0, dx=10.0, dt=1e-3, nsponge=6, freesurface=fs, freqQ=10.0, qMin=0.5,
               qInterior=80.0)
    ocls = {"vti": o2.Oracle2DVTI, "tti": o2.Oracle2DTTI}[kind]
    orc = ocls(nx, nz, 10.0, 8.0, 1e-3, fs=fs, nsponge=6, freqQ=10.0, qMin=0.5, qInterior=80.0)
    arrs = make_model(kind, shape, seed)
    rng = np.random.default_rng(seed + 1)
    for k in ("pCur", "pOld", "mCur", "mOld"):
        arrs[k] = smooth(rng, shape, 1.0, 2)
    for name, a in arrs.items():
        if name in ("sinPhi", "cosPhi"):
            continue
        getattr(w2, _GET[name])(prop)[...] = a.T
        getattr(orc, name)[...] = a
    return prop, orc


ANISO_SHAPES = [(37, 45), (21, 70)]


@pytest.mark.parametrize("kind", ["vti", "tti"])
@pytest.mark.parametrize("fs", [False, True])
@pytest.mark.parametrize("shape", ANISO_SHAPES)
def test_aniso_steps_bit_exact(kind, fs, shape):
    import wavefd_amd.prop2d as w2
    prop, orc = _apair(kind, shape, fs, 21)
    assert np.array_equal(_get(prop, "DtOmegaInvQ"), orc.q)
    for _ in range(6):
        w2.propagateforward(prop)
        orc.time_step()
    for name in ("pCur", "pOld", "mCur", "mOld", "pSpace", "mSpace"):
        assert np.array_equal(_get(prop, _GET[name]), getattr(orc, name)), name


@pytest.mark.parametrize("fs", [False, True])
def test_vti_linear_step_bit_exact(fs):
    import wavefd_amd.prop2d as w2
    prop, orc = _apair("vti", (37, 45), fs, 23)
    for _ in range(4):
        w2.propagateforward_linear(prop)
        orc.time_step_linear()
    for name in ("pCur", "pOld", "mCur", "mOld", "pSpace"):
        assert np.array_equal(_get(prop, _GET[name]), getattr(orc, name)), name


@pytest.mark.parametrize("kind", ["vti", "tti"])
def test_aniso_many_steps(kind):
    import wavefd_amd.prop2d as w2
    prop, orc = _apair(kind, (48, 52), True, 25)
    for _ in range(300):
        w2.propagateforward(prop)
        orc.time_step()
    assert np.array_equal(_get(prop, "PCur"), orc.pCur)
    assert np.array_equal(_get(prop, "MCur"), orc.mCur)


def _amodel(w2, kind, which):
    pre = {"vti": "Prop2DAcoVTIDenQ_DEO2_FDTD", "tti": "Prop2DAcoTTIDenQ_DEO2_FDTD"}[kind]
    return getattr(w2, pre + "_Model_" + which)()


@pytest.mark.parametrize("kind", ["vti", "tti"])
@pytest.mark.parametrize("which", ["V", "VEA"])
@pytest.mark.parametrize("fs", [False, True])
def test_aniso_born_source_bit_exact(kind, which, fs):
    import wavefd_amd.prop2d as w2
    shape = (37, 45)
    prop, orc = _apair(kind, shape, fs, 27)
    rng = np.random.default_rng(28)
    dv, de, da = smooth(rng, shape, 50.0, 2), smooth(rng, shape, 0.02, 2), smooth(rng, shape, 0.02, 2)
    P, M, DP, DM = (smooth(rng, shape, 1.0, 2) for _ in range(4))
    w2.forwardBornInjection(prop, _amodel(w2, kind, which), {"v": dv, "ϵ": de, "η": da},
                            {"pold": P, "mold": M, "pspace": DP, "mspace": DM})
    if which == "V":
        orc.forward_born_injection_V(dv, DP, DM)
    else:
        orc.forward_born_injection_VEA(dv, de, da, P, M, DP, DM)
    assert np.array_equal(_get(prop, "PCur"), orc.pCur)
    assert np.array_equal(_get(prop, "MCur"), orc.mCur)


@pytest.mark.parametrize("kind", ["vti", "tti"])
@pytest.mark.parametrize("which", ["V", "VEA"])
@pytest.mark.parametrize("fs", [False, True])
def test_aniso_imaging_bit_exact(kind, which, fs):
    import wavefd_amd.prop2d as w2
    shape = (37, 45)
    prop, orc = _apair(kind, shape, fs, 29)
    rng = np.random.default_rng(30)
    P, M, DP, DM = (smooth(rng, shape, 1.0, 2) for _ in range(4))
    g = {k: smooth(rng, shape, 1.0, 2) for k in ("v", "ϵ", "η")}
    g2 = {k: a.copy() for k, a in g.items()}
    w2.adjointBornAccumulation(prop, _amodel(w2, kind, which), w2.ImagingConditionStandard(), g,
                               {"pold": P, "mold": M, "pspace": DP, "mspace": DM})
    if which == "V":
        orc.adjoint_born_accumulation_V(g2["v"], DP, DM)
    else:
        orc.adjoint_born_accumulation_VEA(g2["v"], g2["ϵ"], g2["η"], P, M, DP, DM)
    for k in g:
        assert np.array_equal(g[k], g2[k]), k


@pytest.mark.parametrize("kind", ["vti", "tti"])
@pytest.mark.parametrize("fwi", [True, False])
def test_aniso_wavefield_separation(kind, fwi):
    from oracle.oracle import wavefield_separation
    import wavefd_amd.prop2d as w2
    shape = (23, 40)
    prop, orc = _apair(kind, shape, True, 31)
    rng = np.random.default_rng(32)
    DP, DM = smooth(rng, shape, 1.0, 2), smooth(rng, shape, 1.0, 2)
    g = np.zeros(shape, F32)
    ic = w2.ImagingConditionWaveFieldSeparationFWI() if fwi else w2.ImagingConditionWaveFieldSeparationRTM()
    w2.adjointBornAccumulation(prop, _amodel(w2, kind, "V"), ic, {"v": g}, {"pspace": DP, "mspace": DM})
    # float64 restatement of src/prop2DAcoVTIDenQ_DEO2_FDTD.h:509-627 (FFTW's float32 bits are
    # not available): per x column, the nonlinear and adjoint P and M parts summed
    s = (wavefield_separation(DP[:, None, :], orc.pOld[:, None, :], fwi) +
         wavefield_separation(DM[:, None, :], orc.mOld[:, None, :], fwi))[:, 0, :]
    want = 2 * orc.b.astype(np.float64) / orc.v.astype(np.float64) ** 3 * s
    err = np.linalg.norm(g - want) / np.linalg.norm(want)
    assert err <= 2e-5, err


@pytest.mark.parametrize("fs", [False, True])
def test_apply_first_derivatives_host_bit_exact(fs):
    import wavefd_amd.prop2d as w2
    nx, nz = 29, 33
    rng = np.random.default_rng(33)
    a, c = smooth(rng, (nx, nz), 1.0, 2), smooth(rng, (nx, nz), 1.0, 2)
    ox, oz = np.full((nx, nz), 9.0, F32), np.full((nx, nz), 9.0, F32)
    c8 = (1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0)
    idx, idz = F32(1.0 / F32(10.0)), F32(1.0 / F32(8.0))
    w2.applyFirstDerivatives2D_PlusHalf(fs, nx, nz, c8, idx, idz, a, c, ox, oz)
    wx, _ = o2.dplus(o2.Grid2(nx, nz, 10.0, 8.0, fs), a, a)
    _, wz = o2.dplus(o2.Grid2(nx, nz, 10.0, 8.0, fs), c, c)
    assert np.array_equal(ox, wx)
    assert np.array_equal(oz, wz)


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_2000_steps(kind):
    # the parity length the propagators are held to: 2000 steps from a smooth state, identical to
    # the oracle
    import wavefd_amd.prop2d as w2
    if kind == "iso":
        prop, orc = _pair((40, 44), True, 41)
    else:
        prop, orc = _apair(kind, (40, 44), True, 41)
    w2.sync_to_device(prop)  # the host-set state, before the calls stop uploading it
    w2.coherence(prop, "device")
    for _ in range(2000):
        w2.propagateforward(prop)
        orc.time_step()
    w2.sync_to_host(prop)
    assert np.array_equal(_get(prop, "PCur"), orc.pCur)
    if kind != "iso":
        assert np.array_equal(_get(prop, "MCur"), orc.mCur)


FUSED_SHAPES = [(37, 45), (100, 300), (70, 520)]  # (nx, nz): one strip and chunk, several of each


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("fs", [False, True])
@pytest.mark.parametrize("shape", FUSED_SHAPES)
@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_fused_and_two_pass_steps_bit_exact(kind, shape, fs, fused):
    import wavefd_amd.prop2d as w2
    prop, orc = _pair(shape, fs, 51) if kind == "iso" else _apair(kind, shape, fs, 51)
    w2.set_fused_step(prop, fused)
    for _ in range(9):
        w2.propagateforward(prop)
        orc.time_step()
    names = ("pCur", "pOld", "pSpace") + (("mCur", "mOld", "mSpace") if kind != "iso" else ())
    for name in names:
        got = _get(prop, _GET[name]) if kind != "iso" else _get(prop, {"pCur": "PCur", "pOld": "POld",
                                                                        "pSpace": "PSpace"}[name])
        assert np.array_equal(got, getattr(orc, name)), name


@pytest.mark.parametrize("fused", [True, False])
def test_vti_linear_fused_and_two_pass(fused):
    import wavefd_amd.prop2d as w2
    prop, orc = _apair("vti", (100, 300), True, 53)
    w2.set_fused_step(prop, fused)
    for _ in range(5):
        w2.propagateforward_linear(prop)
        orc.time_step_linear()
    for name in ("pCur", "pOld", "mCur", "mOld", "pSpace"):
        assert np.array_equal(_get(prop, _GET[name]), getattr(orc, name)), name


@pytest.mark.parametrize("fs", [False, True])
def test_apply_first_derivatives_host_caller_coefficients(fs):
    # arbitrary c1..c4 and spacings: the entry point must use the caller's, each in its place
    # (src/propagatorStaticFunctions.h:10-161 term order, float32, free-surface odd mirror)
    import wavefd_amd.prop2d as w2
    nx, nz = 23, 30
    rng = np.random.default_rng(61)
    a, c = rng.standard_normal((nx, nz)).astype(F32), rng.standard_normal((nx, nz)).astype(F32)
    cc = [F32(v) for v in (1.3, -0.21, 0.037, -0.0042)]
    idx, idz = F32(0.7), F32(1.9)
    ox, oz = np.full((nx, nz), 9.0, F32), np.full((nx, nz), 9.0, F32)
    w2.applyFirstDerivatives2D_PlusHalf(fs, nx, nz, cc, idx, idz, a, c, ox, oz)

    def sten(f, i, s):  # f[i+s*j] along a stride, j = -3..4
        return (cc[0] * (-f(i) + f(i + s)) + cc[1] * (-f(i - s) + f(i + 2 * s)) + cc[2] * (-f(i - 2 * s) + f(i + 3 * s))
                + cc[3] * (-f(i - 3 * s) + f(i + 4 * s)))
    wx, wz = np.zeros((nx, nz), F32), np.zeros((nx, nz), F32)
    for kx in range(4, nx - 4):
        for kz in range(0 if fs else 4, nz - 4):
            sx = sten(lambda i: a[i, kz], kx, 1)
            sz = sten(lambda i: c[kx, i] if i >= 0 else -c[kx, -i], kz, 1)
            wx[kx, kz] = F32(0) if (fs and kz == 0) else idx * sx
            wz[kx, kz] = idz * sz
    assert np.array_equal(ox, wx)
    assert np.array_equal(oz, wz)


@pytest.mark.parametrize("fs", [False, True])
@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_q_table_and_q_array_give_the_same_step(kind, fs):
    # while DtOmegaInvQ is the SetupDtOmegaInvQ profile the fused step reads Q from a table by
    # sponge distance; once Q's host view or device address is handed out it reads the array:
    # the same bits, the oracle's
    import wavefd_amd.prop2d as w2
    shape = (100, 300)
    mk = (lambda: _pair(shape, fs, 57)) if kind == "iso" else (lambda: _apair(kind, shape, fs, 57))
    (a, orc), (b, _), (c, _) = mk(), mk(), mk()
    _get(b, "DtOmegaInvQ")  # b: the array path from here on
    assert w2.device_ptr(c, "dtOmegaInvQ")  # c: likewise
    for _ in range(7):
        for prop in (a, b, c):
            w2.propagateforward(prop)
        orc.time_step()
    names = ("pCur", "pOld", "pSpace") + (("mCur", "mOld", "mSpace") if kind != "iso" else ())
    for name in names:
        got = [_get(prop, _GET[name]) for prop in (a, b, c)]
        assert np.array_equal(got[0], getattr(orc, name)), name
        assert np.array_equal(got[1], got[0]) and np.array_equal(got[2], got[0]), name


def test_written_q_is_used():
    # Q written through its device address after the profile (a caller's own attenuation model)
    # is what the fused step uses
    import wavefd_amd.prop2d as w2
    from wavefd_amd import ginsu
    shape = (100, 300)
    prop, orc = _pair(shape, True, 59)
    q = (0.002 + 0.01 * np.random.default_rng(5).random(shape)).astype(F32)
    ptr = w2.device_ptr(prop, "dtOmegaInvQ")
    ginsu.copy(ptr, q.ctypes.data, q.nbytes)
    ginsu.sync()
    orc.q[...] = q
    for _ in range(5):
        w2.propagateforward(prop)
        orc.time_step()
    for gname, oname in (("PCur", "pCur"), ("POld", "pOld"), ("PSpace", "pSpace")):
        assert np.array_equal(_get(prop, gname), getattr(orc, oname)), gname
