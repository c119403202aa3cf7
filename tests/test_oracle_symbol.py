"""Pins for the anisotropic terms of the oracle's VTI and TTI steps (CPU only).

The expected values come from the physics, not from the oracle's code:

* the coupled quasi-P / quasi-S system of docs/latex_notes/VtiAcousticQ.tex:52-69
  (b (1+2e) on the horizontal P derivatives, b (1 - f eta^2) on the vertical one,
  b (1-f) on the horizontal M derivatives, b (1 - f + f eta^2) on the vertical one, and the
  coupling b f eta sqrt(1-eta^2) between the vertical derivatives of P and M);
* TTI is that system with the symmetry axis rotated to n = (sin t cos p, sin t sin p, cos t)
  (src/prop3DAcoTTIDenQ_DEO2_FDTD.h:346-355);
* the discrete operators act on a plane wave exp(i k.x) as multiplication by the staggered
  pair's symbol: D+ along axis a gives i K_a exp(i k_a h_a / 2), D- gives i K_a exp(-i k_a h_a / 2),
  with K(k) = (2/h) sum_j c_j sin((2j-1) k h / 2), and c_j the weights that make the pair exact
  on odd monomials (solved here from that condition, not copied).

So one step of a homogeneous medium maps the plane waves P = a_p cos(k.x), M = a_m cos(k.x) to
pSpace = -(S_PP a_p + S_PM a_m) cos(k.x) and mSpace = -(S_MP a_p + S_MM a_m) cos(k.x), with
S_XY = sum_ab K_a K_b cos((k_b h_b - k_a h_a)/2) C^XY_ab for the material matrices C of the
system.  Each test also shows that the symbol with one planted mistake -- (1-f) -> (1+f),
(1+2e) -> (1+e), the coupling's sign, sin p <-> cos p -- misses the oracle by far more than the
tolerance, so the pin is sensitive to exactly those slips.

Energy: with Q = 0 the leapfrog of a symmetric spatial operator conserves
<dP, W dP> + <dM, W dM> (W = b / v^2, dP = P+ - P) / dt^2 - <P+, L_P> - <M+, L_M>; this pins that
the coupling coefficient is the same in both equations and that the rotated sandwich is
symmetric for spatially varying angles.
"""
import numpy as np
import pytest

F32 = np.float32


def _staggered_weights():
    """c_1..c_4 of the 8th-order staggered first derivative: sum_j 2 c_j s_j^(2m+1) = [m == 0]
    for m = 0..3, s_j = j - 1/2 (exact differentiation of x, x^3, x^5, x^7)."""
    s = np.arange(1, 5) - 0.5
    a = np.array([2.0 * s ** (2 * m + 1) for m in range(4)])
    return np.linalg.solve(a, np.array([1.0, 0.0, 0.0, 0.0]))


def _symbol_k(theta, h):
    c = _staggered_weights()
    return (2.0 / h) * sum(c[j] * np.sin((j + 0.5) * theta) for j in range(4))


def _materials(b, eps, eta, f, n, mistake=None):
    """C^PP, C^PM, C^MM (3x3) of the rotated VTI system (VtiAcousticQ.tex:52-69) for the
    symmetry axis n; mistake plants one slip."""
    one_m_f = (1 + f) if mistake == "1+f" else (1 - f)
    e_fac = (1 + eps) if mistake == "1+e" else (1 + 2 * eps)
    coup = b * f * eta * np.sqrt(1 - eta * eta) * (-1 if mistake == "coupling_sign" else 1)
    nn = np.outer(n, n)
    eye = np.eye(3)
    # P: b(1+2e) on the plane normal to n, b(1 - f eta^2) along n
    cpp = b * e_fac * (eye - nn) + b * (1 - f * eta * eta) * nn
    # M: b(1-f) on the plane normal to n, b(1 - f + f eta^2) along n
    cmm = b * one_m_f * (eye - nn) + b * (one_m_f + f * eta * eta) * nn
    cpm = coup * nn
    return cpp, cpm, cmm


def _symbol(theta, h, c):
    kk = np.array([_symbol_k(theta[a], h[a]) for a in range(3)])
    s = 0.0
    for a in range(3):
        for b in range(3):
            s += kk[a] * kk[b] * np.cos((theta[b] - theta[a]) / 2) * c[a, b]
    return s


def _axis(st, ct, sp, cp, mistake=None):
    if mistake == "swap_phi":
        sp, cp = cp, sp
    return np.array([st * cp, st * sp, ct])


SHAPE = (30, 28, 32)
H = (10.0, 12.5, 8.0)
THETA = (0.7, 1.1, 0.45)  # k_a h_a per axis
AP, AM = 1.0, 0.6


def _plane(shape, h, theta):
    ix, iy, iz = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    ph = theta[0] * ix + theta[1] * iy + theta[2] * iz + 0.3
    return np.cos(ph)


def _run(oracle, cls, mat, angles=None):
    prop = cls(*SHAPE, *H, dt=1e-3, fs=False, nsponge=5, qMin=0.0, qInterior=0.0, nthread=4)
    prop.q[...] = 0
    prop.v[...] = 2000.0
    prop.b[...] = mat["b"]
    prop.eps[...] = mat["eps"]
    prop.eta[...] = mat["eta"]
    prop.f[...] = mat["f"]
    if angles is not None:
        prop.sinTheta[...], prop.cosTheta[...], prop.sinPhi[...], prop.cosPhi[...] = angles
    wave = _plane(SHAPE, H, THETA)
    prop.pCur[...] = AP * wave
    prop.mCur[...] = AM * wave
    prop.pOld[...] = prop.pCur
    prop.mOld[...] = prop.mCur
    prop.time_step()
    return prop, wave


def _expected(mat, n, mistake=None):
    cpp, cpm, cmm = _materials(mat["b"], mat["eps"], mat["eta"], mat["f"], n, mistake)
    spp, spm, smm = (_symbol(THETA, H, c) for c in (cpp, cpm, cmm))
    return -(spp * AP + spm * AM), -(spm * AP + smm * AM)


INNER = tuple(slice(8, n - 8) for n in SHAPE)
MISTAKES = ["1+f", "1+e", "coupling_sign"]

VTI_CASES = {
    # eta = 0: M decouples and its operator is b(1-f) times the Laplacian symbol
    "eta0": dict(b=0.8, eps=0.2, eta=0.0, f=0.85),
    "eta": dict(b=0.7, eps=0.25, eta=0.3, f=0.85),
    "f_small": dict(b=1.1, eps=0.1, eta=0.12, f=0.5),
}


def _check(prop, wave, want_p, want_m):
    lp = prop.pSpace.astype(np.float64)[INNER]
    lm = prop.mSpace.astype(np.float64)[INNER]
    w = wave[INNER]
    scale = max(abs(want_p), abs(want_m))
    tol = 2e-5 * scale
    return float(np.abs(lp - want_p * w).max()) / tol, float(np.abs(lm - want_m * w).max()) / tol


@pytest.mark.parametrize("case", sorted(VTI_CASES))
def test_vti_step_matches_plane_wave_symbol(oracle, case):
    mat = VTI_CASES[case]
    prop, wave = _run(oracle, oracle.OracleVTI, mat)
    n = np.array([0.0, 0.0, 1.0])
    ep, em = _check(prop, wave, *_expected(mat, n))
    assert ep <= 1 and em <= 1, (ep, em)
    for mistake in MISTAKES:
        if mistake == "coupling_sign" and mat["eta"] == 0:
            continue  # no coupling at eta = 0
        ep, em = _check(prop, wave, *_expected(mat, n, mistake))
        assert max(ep, em) > 100, (mistake, ep, em)


TTI_CASES = {
    # theta = 0: the VTI axis
    "theta0": (dict(b=0.7, eps=0.25, eta=0.3, f=0.85), 0.0, 0.0),
    "tilted": (dict(b=0.7, eps=0.25, eta=0.3, f=0.85), np.radians(35.0), np.radians(30.0)),
    "tilted_eta0": (dict(b=0.9, eps=0.15, eta=0.0, f=0.85), np.radians(-40.0), np.radians(250.0)),
    "tilted_f": (dict(b=1.2, eps=0.08, eta=0.2, f=0.6), np.radians(20.0), np.radians(115.0)),
}


@pytest.mark.parametrize("case", sorted(TTI_CASES))
def test_tti_step_matches_rotated_plane_wave_symbol(oracle, case):
    mat, t, p = TTI_CASES[case]
    angles = tuple(F32(x) for x in (np.sin(t), np.cos(t), np.sin(p), np.cos(p)))
    prop, wave = _run(oracle, oracle.OracleTTI, mat, angles)
    n = _axis(*[float(a) for a in angles])
    ep, em = _check(prop, wave, *_expected(mat, n))
    assert ep <= 1 and em <= 1, (ep, em)
    mistakes = list(MISTAKES)
    if t != 0.0:
        mistakes.append("swap_phi")
    for mistake in mistakes:
        if mistake == "coupling_sign" and mat["eta"] == 0:
            continue
        n_m = _axis(*[float(a) for a in angles], mistake=mistake)
        ep, em = _check(prop, wave, *_expected(mat, n_m, None if mistake == "swap_phi" else mistake))
        assert max(ep, em) > 100, (mistake, ep, em)


def _random_smooth(rng, shape, amp):
    x = rng.standard_normal(shape)
    for ax in range(3):
        for _ in range(3):
            x = 0.25 * np.roll(x, 1, ax) + 0.5 * x + 0.25 * np.roll(x, -1, ax)
    return (amp * x / np.abs(x).max()).astype(F32)


def _zero_annulus(u):
    u[:4] = u[-4:] = 0
    u[:, :4] = u[:, -4:] = 0
    u[..., :4] = u[..., -4:] = 0
    return u


@pytest.mark.parametrize("physics", ["vti", "tti"])
def test_anisotropic_energy_is_conserved_without_q(oracle, physics):
    rng = np.random.default_rng(11 if physics == "vti" else 12)
    shape = (24, 26, 28)
    cls = oracle.OracleVTI if physics == "vti" else oracle.OracleTTI
    prop = cls(*shape, 10.0, 10.0, 10.0, dt=1e-3, fs=False, qMin=0.0, qInterior=0.0, nthread=4)
    prop.q[...] = 0
    prop.v[...] = 1500 + _random_smooth(rng, shape, 300)
    prop.b[...] = 1.0 + _random_smooth(rng, shape, 0.3)
    prop.eps[...] = 0.15 + _random_smooth(rng, shape, 0.1)
    prop.eta[...] = 0.2 + _random_smooth(rng, shape, 0.1)
    prop.f[...] = 0.85 + _random_smooth(rng, shape, 0.05)
    if physics == "tti":
        th = 0.6 * _random_smooth(rng, shape, 1.0).astype(np.float64)
        ph = 3.0 * _random_smooth(rng, shape, 1.0).astype(np.float64)
        prop.sinTheta[...] = np.sin(th)
        prop.cosTheta[...] = np.cos(th)
        prop.sinPhi[...] = np.sin(ph)
        prop.cosPhi[...] = np.cos(ph)
    for name in ("pCur", "mCur"):
        getattr(prop, name)[...] = _zero_annulus(_random_smooth(rng, shape, 1.0))
    prop.pOld[...] = prop.pCur
    prop.mOld[...] = prop.mCur
    w = prop.b.astype(np.float64) / prop.v.astype(np.float64) ** 2
    dt = 1e-3
    energies = []
    for _ in range(60):
        p0, m0 = prop.pCur.astype(np.float64).copy(), prop.mCur.astype(np.float64).copy()
        prop.time_step()
        p1, m1 = prop.pCur.astype(np.float64), prop.mCur.astype(np.float64)
        kin = np.sum(w * ((p1 - p0) ** 2 + (m1 - m0) ** 2)) / dt ** 2
        pot = -np.sum(p1 * prop.pSpace.astype(np.float64)) - np.sum(m1 * prop.mSpace.astype(np.float64))
        energies.append(kin + pot)
    energies = np.array(energies)
    assert np.abs(energies - energies[0]).max() <= 1e-4 * np.abs(energies[0])
    # the fields moved: the energy is exchanged between its two parts, not trivially constant
    assert np.abs(p1 - p0).max() > 1e-3 * np.abs(p1).max()


def _vti_iso_pair(oracle, f):
    rng = np.random.default_rng(21)
    shape = (22, 20, 24)
    vti = oracle.OracleVTI(*shape, 10.0, 10.0, 10.0, dt=1e-3, fs=True, nsponge=5, freqQ=10.0, qMin=1.0,
                           qInterior=100.0, nthread=4)
    vti.v[...] = 1500 + _random_smooth(rng, shape, 300)
    vti.b[...] = 1.0 + _random_smooth(rng, shape, 0.2)
    vti.f[...] = f
    for name in ("pCur", "pOld", "mCur", "mOld"):
        getattr(vti, name)[...] = _random_smooth(rng, shape, 1.0)
    return vti, shape


def _iso_like(oracle, vti, shape, v, b, cur, old):
    iso = oracle.OracleIso(*shape, 10.0, 10.0, 10.0, dt=1e-3, fs=True, nsponge=5, freqQ=10.0, qMin=1.0,
                           qInterior=100.0, nthread=4)
    iso.v[...] = v
    iso.b[...] = b
    iso.pCur[...] = cur
    iso.pOld[...] = old
    return iso


def test_vti_m_field_reduces_to_iso_when_f_is_zero(oracle):
    """eps = eta = f = 0: M obeys the same equation as P, b times the Laplacian, so both
    fields step exactly as an Iso propagator would (every product with 1 - 0 is exact)."""
    vti, shape = _vti_iso_pair(oracle, 0.0)
    iso_p = _iso_like(oracle, vti, shape, vti.v, vti.b, vti.pCur, vti.pOld)
    iso_m = _iso_like(oracle, vti, shape, vti.v, vti.b, vti.mCur, vti.mOld)
    for _ in range(3):
        vti.time_step()
        iso_p.time_step()
        iso_m.time_step()
    assert np.array_equal(vti.pCur, iso_p.pCur) and np.array_equal(vti.pSpace, iso_p.pSpace)
    assert np.array_equal(vti.mCur, iso_m.pCur) and np.array_equal(vti.mSpace, iso_m.pSpace)


def test_vti_m_field_reduces_to_iso_with_scaled_buoyancy(oracle):
    """eps = eta = 0, f = 0.85: (b/v^2) M_tt = div(b (1-f) grad M), which is an Iso equation for
    buoyancy b(1-f) and velocity v sqrt(1-f) (same b/v^2 weight up to rounding)."""
    f = 0.85
    vti, shape = _vti_iso_pair(oracle, f)
    bm = (vti.b * F32(1 - f)).astype(F32)
    vm = (vti.v * F32(np.sqrt(1 - f))).astype(F32)
    iso_m = _iso_like(oracle, vti, shape, vm, bm, vti.mCur, vti.mOld)
    for _ in range(3):
        vti.time_step()
        iso_m.time_step()
    a, b = vti.mCur, iso_m.pCur
    assert np.allclose(a, b, rtol=0, atol=2e-5 * np.abs(a).max())
    # with the wrong factor (1+f) the M field would not match
    iso_bad = _iso_like(oracle, vti, shape, (vti.v * F32(np.sqrt(1 + f))).astype(F32),
                        (vti.b * F32(1 + f)).astype(F32), vti.mCur, vti.mOld)
    vti2, _ = _vti_iso_pair(oracle, f)
    iso_bad.pCur[...] = vti2.mCur
    iso_bad.pOld[...] = vti2.mOld
    for _ in range(3):
        vti2.time_step()
        iso_bad.time_step()
    assert np.abs(vti2.mCur - iso_bad.pCur).max() > 1e-2 * np.abs(vti2.mCur).max()
