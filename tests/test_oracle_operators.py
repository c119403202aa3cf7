"""Pins for the oracle's derivative operators and step kernels (CPU only).

These checks come from mathematics the reference's scheme fixes, not from re-typing its
formulas: exactness of the 8th-order staggered pair on polynomials, discrete adjointness
of D+ and D- (which is what makes the reference's "self-adjoint" propagator claim hold,
src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:69), closed-form Laplacians of quadratics, and
reductions of TTI to VTI under exact axis rotations.
"""
import numpy as np
import pytest

F32 = np.float32


def _grid(o, n=(20, 21, 22), d=(1.0, 1.0, 1.0), fs=False, dt=1e-3):
    return o.Grid(*n, *d, dt, fs, nthread=4)


def _dot(a, b):
    return float(np.dot(a.astype(np.float64).ravel(), b.astype(np.float64).ravel()))


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("deg", [0, 1, 2, 3, 5, 7])
def test_dplus_dminus_exact_on_polynomials(oracle, axis, deg):
    """Both staggered operators differentiate polynomials of degree <= 7 exactly."""
    g = _grid(oracle, d=(0.5, 0.75, 1.25))
    n = g.shape[axis]
    h = float([g.dx, g.dy, g.dz][axis])
    c = (n - 1) / 2.0
    x = (np.arange(n) - c) * h / 4.0

    def field(xs):
        shape = [1, 1, 1]
        shape[axis] = n
        return np.broadcast_to(np.reshape(xs, shape), g.shape).astype(F32).copy()

    f = field(x ** deg)
    dp = oracle.dplus(g, f, f, f)[axis]
    dm = oracle.dminus(g, f, f, f)[axis]
    # D+ at k is the derivative at k+1/2 (x advances h/4 per cell)
    xp = x + h / 8.0
    xm = x - h / 8.0
    expect_p = field(deg * xp ** (deg - 1) / 4.0 if deg > 0 else 0 * xp)
    expect_m = field(deg * xm ** (deg - 1) / 4.0 if deg > 0 else 0 * xm)
    sl = [slice(4, s - 4) for s in g.shape]
    scale = max(1.0, float(np.abs(expect_p[tuple(sl)]).max()))
    assert np.allclose(dp[tuple(sl)], expect_p[tuple(sl)], atol=2e-5 * scale, rtol=0)
    assert np.allclose(dm[tuple(sl)], expect_m[tuple(sl)], atol=2e-5 * scale, rtol=0)
    # the 4-cell annulus is zeroed
    assert np.all(dp[:4] == 0) and np.all(dp[-4:] == 0)
    assert np.all(dm[:, :4] == 0) and np.all(dm[..., -4:] == 0)


def _interior_support(g, rng, fs, surface_zero=False):
    """Random field on the update region; with a free surface the region reaches kz=0.
    surface_zero sets the kz=0 row to 0 (p = 0 at the surface), the condition under which
    the odd (D+) and even (D-) mirrors are exact transposes."""
    u = rng.standard_normal(g.shape).astype(F32)
    mask = np.zeros(g.shape, bool)
    z0 = 0 if fs else 4
    mask[4:-4, 4:-4, z0:-4] = True
    u[~mask] = 0
    if surface_zero:
        u[..., 0] = 0
    return u


@pytest.mark.parametrize("fs", [False, True])
def test_dplus_is_minus_transpose_of_dminus(oracle, fs):
    """<D+ u, w> = -<u, D- w> on fields supported where the step updates."""
    rng = np.random.default_rng(11)
    g = _grid(oracle, n=(17, 18, 19), d=(10.0, 12.5, 7.5), fs=fs)
    u = _interior_support(g, rng, fs, surface_zero=True)
    w = [_interior_support(g, rng, fs) for _ in range(3)]
    dp = oracle.dplus(g, u, u, u)
    dm = oracle.dminus(g, *w)
    for a in range(3):
        lhs = _dot(dp[a], w[a])
        rhs = -_dot(u, dm[a])
        # normalise by the size of the summed terms: float32 outputs, float64 sums
        scale = _dot(np.abs(dp[a]), np.abs(w[a]))
        assert abs(lhs - rhs) <= 1e-6 * scale, (a, lhs, rhs, scale)


@pytest.mark.parametrize("fs", [False, True])
def test_rotated_pair_is_adjoint(oracle, fs):
    """The TTI rotated operators R+ and R- are minus-transposes of each other."""
    rng = np.random.default_rng(5)
    g = _grid(oracle, n=(17, 16, 18), d=(10.0, 12.0, 8.0), fs=fs)
    th = rng.uniform(-0.8, 0.8, g.shape).astype(F32)
    ph = rng.uniform(0, 6.28, g.shape).astype(F32)
    ang = (np.sin(th), np.cos(th), np.sin(ph), np.cos(ph))
    u = _interior_support(g, rng, fs, surface_zero=True)
    w = [_interior_support(g, rng, fs) for _ in range(3)]
    rp = oracle.tti_dplus_rot(g, u, u, u, *ang)
    lhs = sum(_dot(rp[a], w[a]) for a in range(3))
    rm = oracle.tti_dminus_rot(g, *w, *ang)
    rhs = -_dot(u, rm[0] + rm[1] + rm[2])
    scale = sum(_dot(np.abs(rp[a]), np.abs(w[a])) for a in range(3))
    assert abs(lhs - rhs) <= 1e-6 * scale, (lhs, rhs, scale)


def test_rotated_reduces_to_cartesian_at_theta_zero(oracle):
    rng = np.random.default_rng(3)
    g = _grid(oracle, fs=True)
    one = np.ones(g.shape, F32)
    zero = np.zeros(g.shape, F32)
    u = rng.standard_normal(g.shape).astype(F32)
    # theta=0, phi=0: g1 = dx, g2 = dy, g3 = dz
    rp = oracle.tti_dplus_rot(g, u, u, u, zero, one, zero, one)
    dp = oracle.dplus(g, u, u, u)
    for a in range(3):
        assert np.allclose(rp[a], dp[a], rtol=0, atol=1e-6 * np.abs(dp[a]).max())
    w = [rng.standard_normal(g.shape).astype(F32) for _ in range(3)]
    rm = oracle.tti_dminus_rot(g, *w, zero, one, zero, one)
    dm = oracle.dminus(g, *w)
    for a in range(3):
        assert np.allclose(rm[a], dm[a], rtol=0, atol=1e-6 * np.abs(dm[a]).max())


def test_iso_step_laplacian_of_quadratic(oracle):
    """p = a x^2 + b y^2 + c z^2 gives L p = 2 B (a+b+c) away from the annulus."""
    nx, ny, nz = 24, 22, 26
    d = (10.0, 12.5, 8.0)
    prop = oracle.OracleIso(nx, ny, nz, *d, dt=1e-3, fs=False, qMin=0.0, qInterior=0.0, nthread=4)
    X, Y, Z = np.meshgrid(np.arange(nx) * d[0], np.arange(ny) * d[1], np.arange(nz) * d[2], indexing="ij")
    a, b_, c = 3e-4, -2e-4, 5e-4
    prop.pCur[...] = (a * (X - 100) ** 2 + b_ * (Y - 120) ** 2 + c * (Z - 90) ** 2).astype(F32)
    prop.pOld[...] = prop.pCur
    B = F32(0.7)
    prop.b[...] = B
    prop.v[...] = 2000.0
    p0 = prop.pCur.copy()
    prop.time_step()
    L = prop.pSpace[8:-8, 8:-8, 8:-8]
    assert np.allclose(L, 2 * B * (a + b_ + c), rtol=2e-3, atol=0)
    # update with Q=0 and pOld = pCur: p+ = p + dt^2 v^2/b L
    expect = p0 + (1e-3) ** 2 * 2000.0 ** 2 / B * 2 * B * (a + b_ + c)
    assert np.allclose(prop.pCur[8:-8, 8:-8, 8:-8], expect[8:-8, 8:-8, 8:-8], rtol=1e-5)


def _random_smooth(rng, shape, amp):
    x = rng.standard_normal(shape)
    for ax in range(3):
        for _ in range(3):
            x = 0.25 * np.roll(x, 1, ax) + 0.5 * x + 0.25 * np.roll(x, -1, ax)
    return (amp * x / np.abs(x).max()).astype(F32)


def test_iso_energy_is_conserved_without_q(oracle):
    """Leapfrog with symmetric L conserves <dp, M dp>/dt^2 - <p+, L p> (M = b/v^2)."""
    rng = np.random.default_rng(7)
    nx, ny, nz = 26, 24, 28
    prop = oracle.OracleIso(nx, ny, nz, 10.0, 10.0, 10.0, dt=1e-3, fs=False, qMin=0.0, qInterior=0.0, nthread=4)
    prop.v[...] = 1500 + _random_smooth(rng, prop.v.shape, 500)
    prop.b[...] = 1.0 + _random_smooth(rng, prop.b.shape, 0.3)
    p = _random_smooth(rng, prop.v.shape, 1.0)
    p[:4] = p[-4:] = 0
    p[:, :4] = p[:, -4:] = 0
    p[..., :4] = p[..., -4:] = 0
    prop.pCur[...] = p
    prop.pOld[...] = p
    M = (prop.b.astype(np.float64) / prop.v.astype(np.float64) ** 2)
    dt = 1e-3
    energies = []
    for _ in range(60):
        prev = prop.pCur.astype(np.float64).copy()
        prop.time_step()
        new = prop.pCur.astype(np.float64)
        e = np.sum(M * (new - prev) ** 2) / dt ** 2 - np.sum(new * prop.pSpace.astype(np.float64))
        energies.append(e)
    energies = np.array(energies)
    assert np.abs(energies - energies[0]).max() <= 1e-4 * np.abs(energies[0])


def _vti_like(o, cls, shape, d, fs, rng):
    prop = cls(*shape, *d, dt=1e-3, fs=fs, nsponge=5, freqQ=10.0, qMin=1.0, qInterior=100.0, nthread=4)
    prop.v[...] = 1500 + _random_smooth(rng, shape, 300)
    prop.b[...] = 1.0 + _random_smooth(rng, shape, 0.2)
    prop.eps[...] = 0.1 + _random_smooth(rng, shape, 0.05)
    prop.eta[...] = 0.05 + _random_smooth(rng, shape, 0.03)
    prop.f[...] = 0.85
    for name in ("pCur", "pOld", "mCur", "mOld"):
        getattr(prop, name)[...] = _random_smooth(rng, shape, 1.0)
    return prop


def test_vti_reduces_to_iso_when_isotropic(oracle):
    rng = np.random.default_rng(2)
    shape = (20, 19, 21)
    vti = _vti_like(oracle, oracle.OracleVTI, shape, (10.0, 10.0, 10.0), True, rng)
    vti.eps[...] = 0
    vti.eta[...] = 0
    iso = oracle.OracleIso(*shape, 10.0, 10.0, 10.0, dt=1e-3, fs=True, nsponge=5, freqQ=10.0, qMin=1.0,
                           qInterior=100.0, nthread=4)
    iso.v[...] = vti.v
    iso.b[...] = vti.b
    iso.pCur[...] = vti.pCur
    iso.pOld[...] = vti.pOld
    for _ in range(3):
        vti.time_step()
        iso.time_step()
    assert np.array_equal(vti.pCur, iso.pCur)
    assert np.array_equal(vti.pSpace, iso.pSpace)


def test_tti_reduces_to_vti_at_theta_zero(oracle):
    rng = np.random.default_rng(4)
    shape = (20, 19, 21)
    vti = _vti_like(oracle, oracle.OracleVTI, shape, (10.0, 12.0, 8.0), True, rng)
    tti = oracle.OracleTTI(*shape, 10.0, 12.0, 8.0, dt=1e-3, fs=True, nsponge=5, freqQ=10.0, qMin=1.0,
                           qInterior=100.0, nthread=4)
    for name in ("v", "b", "eps", "eta", "f", "pCur", "pOld", "mCur", "mOld"):
        getattr(tti, name)[...] = getattr(vti, name)
    tti.cosTheta[...] = 1
    tti.cosPhi[...] = 1
    for _ in range(4):
        vti.time_step()
        tti.time_step()
    for name in ("pCur", "mCur", "pSpace", "mSpace"):
        a, b = getattr(vti, name), getattr(tti, name)
        assert np.allclose(a, b, rtol=0, atol=2e-5 * np.abs(a).max()), name


@pytest.mark.parametrize("axis", [0, 1])
def test_tti_axis_rotation_matches_transposed_vti(oracle, axis):
    """theta=90deg puts the symmetry axis along x (phi=0) or y (phi=90deg); that is VTI on a
    grid whose z axis is swapped with x or y.  Pins n = (sin t cos p, sin t sin p, cos t)."""
    rng = np.random.default_rng(9)
    n = 20
    shape = (n, n, n)
    tti = _vti_like(oracle, oracle.OracleTTI, shape, (10.0, 10.0, 10.0), False, rng)
    tti.q[...] = 0
    tti.sinTheta[...] = 1
    tti.cosTheta[...] = 0
    if axis == 0:
        tti.cosPhi[...] = 1
        perm = (2, 1, 0)
    else:
        tti.sinPhi[...] = 1
        perm = (0, 2, 1)
    vti = oracle.OracleVTI(*shape, 10.0, 10.0, 10.0, dt=1e-3, fs=False, nsponge=5, qMin=0.0, qInterior=0.0,
                           nthread=4)
    for name in ("v", "b", "eps", "eta", "f", "pCur", "pOld", "mCur", "mOld"):
        getattr(vti, name)[...] = np.ascontiguousarray(getattr(tti, name).transpose(perm))
    for _ in range(4):
        tti.time_step()
        vti.time_step()
    for name in ("pCur", "mCur"):
        a = getattr(tti, name)
        b = np.ascontiguousarray(getattr(vti, name).transpose(perm))
        assert np.allclose(a, b, rtol=0, atol=2e-5 * np.abs(a).max()), name
