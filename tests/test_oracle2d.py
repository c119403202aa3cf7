"""Pins for the 2D Iso oracle (oracle/oracle2d.py, CPU only), from mathematics the scheme fixes
and from the pinned 3D oracle:

* D+ / D- differentiate polynomials of degree <= 7 exactly along x and z, zero annulus;
* <D+ u, w> = -<u, D- w> with and without the free-surface mirrors;
* the 2D Iso step equals the 3D oracle's step from the same state extended constantly in y
  (the y-derivatives vanish, L = dx + (0) + dz), in the y-row the 3D update covers, same Q;
* the 2D Q profile: qMin at the edges, qInterior inside, the free-surface top not sponged;
* each Born source is the derivative of the step in the model (Richardson-extrapolated central
  differences), each imaging condition its adjoint in the b/v^2-weighted inner product.
"""
import numpy as np
import pytest

from oracle import oracle2d as o2
from tests.helpers import smooth

F32 = np.float32
SHAPE = (26, 30)  # (nx, nz)


def _dot(a, b):
    return float(np.dot(a.astype(np.float64).ravel(), b.astype(np.float64).ravel()))


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("deg", [0, 1, 2, 3, 5, 7])
def test_dplus_dminus_exact_on_polynomials(axis, deg):
    g = o2.Grid2(*SHAPE, 0.5, 1.25, False)
    n = SHAPE[axis]
    h = [0.5, 1.25][axis]
    x = (np.arange(n) - (n - 1) / 2.0) * h / 4.0

    def field(xs):
        shape = [1, 1]
        shape[axis] = n
        return np.broadcast_to(np.reshape(xs, shape), SHAPE).astype(F32).copy()

    f = field(x ** deg)
    dp = o2.dplus(g, f, f)[axis]
    dm = o2.dminus(g, f, f)[axis]
    ep = field(deg * (x + h / 8.0) ** (deg - 1) / 4.0 if deg > 0 else 0 * x)
    em = field(deg * (x - h / 8.0) ** (deg - 1) / 4.0 if deg > 0 else 0 * x)
    sl = (slice(4, -4), slice(4, -4))
    scale = max(1.0, float(np.abs(ep[sl]).max()))
    assert np.allclose(dp[sl], ep[sl], atol=2e-5 * scale, rtol=0)
    assert np.allclose(dm[sl], em[sl], atol=2e-5 * scale, rtol=0)
    assert np.all(dp[:4] == 0) and np.all(dp[-4:] == 0) and np.all(dm[:, -4:] == 0)


def _support(rng, fs, surface_zero=False):
    u = rng.standard_normal(SHAPE).astype(F32)
    m = np.zeros(SHAPE, bool)
    m[4:-4, (0 if fs else 4):-4] = True
    u[~m] = 0
    if surface_zero:
        u[:, 0] = 0
    return u


@pytest.mark.parametrize("fs", [False, True])
def test_dplus_is_minus_transpose_of_dminus(fs):
    rng = np.random.default_rng(3)
    g = o2.Grid2(*SHAPE, 10.0, 7.5, fs)
    u = _support(rng, fs, surface_zero=True)
    w = [_support(rng, fs) for _ in range(2)]
    dp = o2.dplus(g, u, u)
    dm = o2.dminus(g, *w)
    for a in range(2):
        lhs, rhs = _dot(dp[a], w[a]), -_dot(u, dm[a])
        assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs)), (a, lhs, rhs)


@pytest.mark.parametrize("fs", [False, True])
def test_step_equals_3d_step_on_y_constant_fields(oracle, fs):
    rng = np.random.default_rng(5)
    nx, nz = SHAPE
    ny = 9
    p2 = o2.Oracle2DIso(nx, nz, 10.0, 8.0, 1e-3, fs=fs, nsponge=6, freqQ=10.0, qMin=0.5, qInterior=80.0)
    p3 = oracle.OracleIso(nx, ny, nz, 10.0, 12.0, 8.0, dt=1e-3, fs=fs, nsponge=6, freqQ=10.0, qMin=0.5,
                          qInterior=80.0, nthread=4)
    v = (2000 + smooth(rng, SHAPE, 300.0, 2)).astype(F32)
    b = (0.8 + smooth(rng, SHAPE, 0.2, 2)).astype(F32)
    pc = smooth(rng, SHAPE, 1.0, 2)
    po = smooth(rng, SHAPE, 1.0, 2)
    for name, a in (("v", v), ("b", b), ("pCur", pc), ("pOld", po)):
        getattr(p2, name)[...] = a
    p3.q[...] = p2.q[:, None, :]
    for step in range(3):  # each 2D step against one 3D step from the same (y-constant) state
        for name in ("v", "b", "pCur", "pOld"):
            getattr(p3, name)[...] = getattr(p2, name)[:, None, :]
        p2.time_step()
        p3.time_step()
        for name in ("pCur", "pOld", "pSpace"):
            assert np.array_equal(getattr(p2, name), getattr(p3, name)[:, 4, :]), (step, name)


def test_q_profile():
    nx, nz, ns = 40, 36, 8
    dt, fq, qmin, qint = 1e-3, 10.0, 0.5, 80.0
    q = o2.setup_q(True, nx, nz, ns, dt, fq, qmin, qint)
    want_int = F32(dt * 2.0 * 3.1415926535897 * fq / qint)
    want_edge = F32(dt * 2.0 * 3.1415926535897 * fq / F32(np.exp(F32(np.log(qmin)))))
    assert q[20, 5] == want_int  # the free-surface top is not sponged
    assert q[0, 20] == want_edge and q[20, nz - 1] == want_edge
    assert np.all(np.diff(q[:ns, 20]) < 0)  # the ramp towards the interior
    q2 = o2.setup_q(False, nx, nz, ns, dt, fq, qmin, qint)
    assert q2[20, 0] == want_edge
    z = o2.setup_q(True, nx, nz, ns, dt, fq, 0.0, 0.0)  # unphysical: the reference clears 2 values
    assert z[0, 0] == 0 and z[0, 1] == 0


def _interior(a, fs):
    a = a.copy()
    a[:4] = a[-4:] = 0
    a[:, -4:] = 0
    if fs:
        a[:, 0] = 0
    else:
        a[:, :4] = 0
    return a


def _model(seed):
    rng = np.random.default_rng(seed)
    return {"v": (2000 + smooth(rng, SHAPE, 300.0, 2)).astype(F32), "b": (0.8 + smooth(rng, SHAPE, 0.2, 2)).astype(F32)}


def _prop(fs, model):
    p = o2.Oracle2DIso(*SHAPE, 10.0, 9.0, 1e-3, fs=fs, nsponge=4, qMin=0.0, qInterior=0.0)
    for k, v in model.items():
        getattr(p, k)[...] = v
    return p


def _step(fs, model, fields):
    p = _prop(fs, model)
    for k, v in fields.items():
        getattr(p, k)[...] = v
    p.time_step()
    return p


@pytest.mark.parametrize("which", ["V", "VB", "B"])
@pytest.mark.parametrize("fs", [False, True])
def test_born_source_is_the_linearised_step(which, fs):
    rng = np.random.default_rng(17)
    model = _model(23)
    fields = {"pCur": _interior(smooth(rng, SHAPE, 1.0, 2), fs), "pOld": _interior(smooth(rng, SHAPE, 1.0, 2), fs)}
    dm = {}
    if "V" in which:
        dm["v"] = smooth(rng, SHAPE, 100.0, 2)
    if "B" in which:
        dm["b"] = smooth(rng, SHAPE, 0.02, 2)

    def fd(h):
        mp = {k: (v + (h * dm[k] if k in dm else 0)).astype(F32) for k, v in model.items()}
        mm = {k: (v - (h * dm[k] if k in dm else 0)).astype(F32) for k, v in model.items()}
        return (_step(fs, mp, fields).pCur.astype(np.float64) - _step(fs, mm, fields).pCur) / (2 * h)

    want = (4 * fd(0.5) - fd(1.0)) / 3
    base = _step(fs, model, fields)
    base.scale_spatial_derivatives()
    lin = _prop(fs, model)
    P, DP = fields["pCur"], base.pSpace
    if which == "V":
        lin.forward_born_injection_V(dm["v"], DP)
    elif which == "VB":
        lin.forward_born_injection_VB(dm["v"], dm["b"], P, DP)
    else:
        lin.forward_born_injection_B(dm["b"], P, DP)
    got = lin.pCur.astype(np.float64)
    assert np.linalg.norm(got - want) <= 1e-3 * np.linalg.norm(want)


@pytest.mark.parametrize("which", ["V", "VB", "B"])
@pytest.mark.parametrize("fs", [False, True])
def test_imaging_condition_is_adjoint_of_source(which, fs):
    rng = np.random.default_rng(31)
    model = _model(37)
    fields = {"pCur": _interior(smooth(rng, SHAPE, 1.0, 2), fs), "pOld": _interior(smooth(rng, SHAPE, 1.0, 2), fs)}
    base = _step(fs, model, fields)
    base.scale_spatial_derivatives()
    P, DP = fields["pCur"], base.pSpace.copy()
    lam = _interior(smooth(rng, SHAPE, 1.0, 2), fs)
    names = {"V": ["v"], "VB": ["v", "b"], "B": ["b"]}[which]
    dm = {k: _interior(smooth(rng, SHAPE, 1.0, 2), fs) for k in names}
    lin, img = _prop(fs, model), _prop(fs, model)
    img.pOld[...] = lam
    g = {k: np.zeros(SHAPE, F32) for k in names}
    if which == "V":
        lin.forward_born_injection_V(dm["v"], DP)
        img.adjoint_born_accumulation_V(g["v"], DP)
    elif which == "VB":
        lin.forward_born_injection_VB(dm["v"], dm["b"], P, DP)
        img.adjoint_born_accumulation_VB(g["v"], g["b"], P, DP)
    else:
        lin.forward_born_injection_B(dm["b"], P, DP)
        img.adjoint_born_accumulation_B(g["b"], P, DP)
    lhs = np.sum(lin.pCur.astype(np.float64) * lam * (model["b"] / model["v"].astype(np.float64) ** 2))
    rhs = 1e-3 ** 2 * sum(np.sum(dm[k].astype(np.float64) * g[k]) for k in names)
    assert abs(lhs - rhs) <= 1e-4 * (abs(lhs) + abs(rhs)), (lhs, rhs)
