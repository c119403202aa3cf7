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


# ---- 2D VTI / TTI ---------------------------------------------------------------------------

def _aniso(kind, fs, seed, shape=SHAPE, dx=10.0, dz=8.0, q=True):
    from tests.helpers import make_model
    cls = {"iso": o2.Oracle2DIso, "vti": o2.Oracle2DVTI, "tti": o2.Oracle2DTTI}[kind]
    kw = dict(fs=fs, nsponge=6, freqQ=10.0, qMin=0.5, qInterior=80.0) if q else dict(fs=fs, nsponge=4, qMin=0.0,
                                                                                      qInterior=0.0)
    p = cls(*shape, dx, dz, 1e-3, **kw)
    for k, v in make_model(kind, shape, seed).items():
        if hasattr(p, k):
            getattr(p, k)[...] = v
    rng = np.random.default_rng(seed + 1)
    for k in ("pCur", "pOld", "mCur", "mOld"):
        if hasattr(p, k):
            getattr(p, k)[...] = smooth(rng, shape, 1.0, 2)
    return p


@pytest.mark.parametrize("fs", [False, True])
def test_vti_without_anisotropy_is_iso(fs):
    vti = _aniso("vti", fs, 3)
    vti.eps[...] = 0
    vti.eta[...] = 0
    iso = o2.Oracle2DIso(*SHAPE, 10.0, 8.0, 1e-3, fs=fs, nsponge=6, freqQ=10.0, qMin=0.5, qInterior=80.0)
    for k in ("v", "b", "pCur", "pOld"):
        getattr(iso, k)[...] = getattr(vti, k)
    for _ in range(4):
        vti.time_step()
        iso.time_step()
    for k in ("pCur", "pOld", "pSpace"):
        assert np.array_equal(getattr(vti, k), getattr(iso, k)), k


@pytest.mark.parametrize("eta_zero", [True, False])
@pytest.mark.parametrize("fs", [False, True])
def test_vti_step_equals_3d_step_on_y_constant_fields(oracle, fs, eta_zero):
    nx, nz = SHAPE
    ny = 9
    p2 = _aniso("vti", fs, 5)
    if eta_zero:
        p2.eta[...] = 0
    p3 = oracle.OracleVTI(nx, ny, nz, 10.0, 12.0, 8.0, dt=1e-3, fs=fs, nsponge=6, freqQ=10.0, qMin=0.5,
                          qInterior=80.0, nthread=4)
    p3.q[...] = p2.q[:, None, :]
    names = ("v", "b", "eps", "eta", "f", "pCur", "pOld", "mCur", "mOld")
    for step in range(3):
        for name in names:
            getattr(p3, name)[...] = getattr(p2, name)[:, None, :]
        p2.time_step()
        p3.time_step()
        for name in ("pCur", "pOld", "mCur", "mOld", "pSpace", "mSpace"):
            a, b = getattr(p2, name), getattr(p3, name)[:, 4, :]
            if eta_zero:
                assert np.array_equal(a, b), (step, name)
            else:
                assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max(), (step, name)


@pytest.mark.parametrize("fs", [False, True])
def test_vti_linear_step(fs):
    a, b = _aniso("vti", fs, 7), _aniso("vti", fs, 7)
    b.pSpace[...] = 7.0
    a.time_step()
    b.time_step_linear()
    for k in ("pCur", "pOld", "mCur", "mOld"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert np.all(b.pSpace == 7.0) and not np.all(a.pSpace == 7.0)


def _angles(p, theta):
    p.sinTheta[...] = F32(np.sin(theta)) if theta not in (0.0, 90.0) else (0.0 if theta == 0.0 else 1.0)
    p.cosTheta[...] = F32(np.cos(theta)) if theta not in (0.0, 90.0) else (1.0 if theta == 0.0 else 0.0)


@pytest.mark.parametrize("fs", [False, True])
def test_tti_at_theta_zero_is_vti(fs):
    tti, vti = _aniso("tti", fs, 9), _aniso("vti", fs, 9)
    _angles(tti, 0.0)
    for _ in range(4):
        tti.time_step()
        vti.time_step()
    for k in ("pCur", "pOld", "mCur", "mOld", "pSpace", "mSpace"):
        assert np.array_equal(getattr(tti, k), getattr(vti, k)), k


def test_tti_at_theta_90_is_vti_on_the_transposed_grid():
    # symmetry axis along x: the VTI scheme with x and z exchanged (no free surface, which
    # would break the symmetry), dx and dz exchanged with them
    tti = _aniso("tti", False, 11, dx=10.0, dz=8.0)
    _angles(tti, 90.0)
    nx, nz = SHAPE
    vti = o2.Oracle2DVTI(nz, nx, 8.0, 10.0, 1e-3, fs=False, nsponge=6, freqQ=10.0, qMin=0.5, qInterior=80.0)
    for k in ("v", "b", "eps", "eta", "f", "q", "pCur", "pOld", "mCur", "mOld"):
        getattr(vti, k)[...] = getattr(tti, k).T
    for _ in range(4):
        tti.time_step()
        vti.time_step()
    for k in ("pCur", "pOld", "mCur", "mOld", "pSpace", "mSpace"):
        assert np.array_equal(getattr(tti, k), getattr(vti, k).T), k


@pytest.mark.parametrize("fs", [False, True])
def test_tti_rotated_dplus_is_minus_transpose_of_rotated_dminus(fs):
    # <R+ u, w> = <g1, w1> + <g3, w3> = -<u, R-(w)_1 + R-(w)_3>: the time update's sum
    rng = np.random.default_rng(13)
    g = o2.Grid2(*SHAPE, 10.0, 7.5, fs)
    th = smooth(rng, SHAPE, 1.2, 2)
    sT, cT = np.sin(th).astype(F32), np.cos(th).astype(F32)
    u = _support(rng, fs, surface_zero=True)
    w1, w3 = _support(rng, fs), _support(rng, fs)
    g1, g3 = o2.tti_dplus(g, u, u, sT, cT)
    o1, o3 = o2.tti_dminus(g, w1, w3, sT, cT)
    lhs = _dot(g1, w1) + _dot(g3, w3)
    rhs = -_dot(u, o1.astype(np.float64) + o3)
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs)), (lhs, rhs)


ANISO = [("vti", "V"), ("vti", "VEA"), ("tti", "V"), ("tti", "VEA")]


def _aniso_model(kind, seed):
    from tests.helpers import make_model
    m = make_model(kind, SHAPE, seed)
    m.pop("sinPhi", None)
    m.pop("cosPhi", None)
    return m


def _aprop(kind, fs, model):
    cls = {"vti": o2.Oracle2DVTI, "tti": o2.Oracle2DTTI}[kind]
    p = cls(*SHAPE, 10.0, 9.0, 1e-3, fs=fs, nsponge=4, qMin=0.0, qInterior=0.0)
    for k, v in model.items():
        getattr(p, k)[...] = v
    return p


def _astep(kind, fs, model, fields):
    p = _aprop(kind, fs, model)
    for k, v in fields.items():
        getattr(p, k)[...] = v
    p.time_step()
    return p


def _afields(fs, seed):
    rng = np.random.default_rng(seed)
    return {k: _interior(smooth(rng, SHAPE, 1.0, 2), fs) for k in ("pCur", "pOld", "mCur", "mOld")}


def _ainject(lin, which, dm, P, M, DP, DM):
    if which == "V":
        lin.forward_born_injection_V(dm["v"], DP, DM)
    else:
        lin.forward_born_injection_VEA(dm["v"], dm["eps"], dm["eta"], P, M, DP, DM)


@pytest.mark.parametrize("kind,which", ANISO)
@pytest.mark.parametrize("fs", [False, True])
def test_aniso_born_source_is_the_linearised_step(kind, which, fs):
    rng = np.random.default_rng(19)
    model = _aniso_model(kind, 23)
    fields = _afields(fs, 29)
    dm = {"v": smooth(rng, SHAPE, 100.0, 2)}
    if which == "VEA":
        dm["eps"] = smooth(rng, SHAPE, 0.02, 2)
        dm["eta"] = smooth(rng, SHAPE, 0.02, 2)

    def fd(h):
        mp = {k: (v + (h * dm[k] if k in dm else 0)).astype(F32) for k, v in model.items()}
        mm = {k: (v - (h * dm[k] if k in dm else 0)).astype(F32) for k, v in model.items()}
        sp, sm = _astep(kind, fs, mp, fields), _astep(kind, fs, mm, fields)
        return [(getattr(sp, k).astype(np.float64) - getattr(sm, k)) / (2 * h) for k in ("pCur", "mCur")]

    f1, f2 = fd(1.0), fd(0.5)
    want = [(4 * b - a) / 3 for a, b in zip(f1, f2)]
    base = _astep(kind, fs, model, fields)
    base.scale_spatial_derivatives()
    lin = _aprop(kind, fs, model)
    _ainject(lin, which, dm, fields["pCur"], fields["mCur"], base.pSpace, base.mSpace)
    for got, w in zip((lin.pCur, lin.mCur), want):
        assert np.linalg.norm(got - w) <= 1e-3 * np.linalg.norm(w)


@pytest.mark.parametrize("kind,which", ANISO)
@pytest.mark.parametrize("fs", [False, True])
def test_aniso_imaging_condition_is_adjoint_of_source(kind, which, fs):
    rng = np.random.default_rng(31)
    model = _aniso_model(kind, 37)
    fields = _afields(fs, 41)
    base = _astep(kind, fs, model, fields)
    base.scale_spatial_derivatives()
    P, M, DP, DM = fields["pCur"], fields["mCur"], base.pSpace.copy(), base.mSpace.copy()
    lam_p, lam_m = _interior(smooth(rng, SHAPE, 1.0, 2), fs), _interior(smooth(rng, SHAPE, 1.0, 2), fs)
    names = ["v"] if which == "V" else ["v", "eps", "eta"]
    dm = {k: _interior(smooth(rng, SHAPE, 1.0, 2), fs) for k in names}
    lin, img = _aprop(kind, fs, model), _aprop(kind, fs, model)
    img.pOld[...] = lam_p
    img.mOld[...] = lam_m
    g = {k: np.zeros(SHAPE, F32) for k in names}
    _ainject(lin, which, dm, P, M, DP, DM)
    if which == "V":
        img.adjoint_born_accumulation_V(g["v"], DP, DM)
    else:
        img.adjoint_born_accumulation_VEA(g["v"], g["eps"], g["eta"], P, M, DP, DM)
    w_ = model["b"] / model["v"].astype(np.float64) ** 2
    lhs = np.sum((lin.pCur.astype(np.float64) * lam_p + lin.mCur.astype(np.float64) * lam_m) * w_)
    rhs = 1e-3 ** 2 * sum(np.sum(dm[k].astype(np.float64) * g[k]) for k in names)
    assert abs(lhs - rhs) <= 1e-4 * (abs(lhs) + abs(rhs)), (lhs, rhs)
