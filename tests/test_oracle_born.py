"""Pins for the oracle's Born sources and imaging conditions (CPU only).

1. Linearisation: each Born source is the derivative of the (pinned) nonlinear step with
   respect to the model, applied to the perturbation.  A central finite difference of the
   step in the model, (S(m + h dm) - S(m - h dm)) / 2h, must equal the source the oracle
   injects given DP = (v^2/b) L p from the unperturbed step (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:
   293-300 and :500-513 describe that contract).
2. Duality: each imaging condition is the adjoint of its source in the b/v^2-weighted inner
   product (the weighting that makes the propagator "self-adjoint",
   src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:69): <src(dm), lam * b/v^2> = dt^2 <dm, img(lam)>.
A dropped term, a wrong factor or sign, or a transposed operand fails one of these.
"""
import numpy as np
import pytest

from tests.helpers import make_model, smooth

F32 = np.float32
SHAPE = (22, 21, 24)


def _interior(arr, fs):
    a = arr.copy()
    a[:4] = a[-4:] = 0
    a[:, :4] = a[:, -4:] = 0
    a[..., -4:] = 0
    if fs:
        a[..., 0] = 0
    else:
        a[..., :4] = 0
    return a


def _prop(o, kind, fs, model, q=False):
    cls = {"iso": o.OracleIso, "vti": o.OracleVTI, "tti": o.OracleTTI}[kind]
    p = cls(*SHAPE, 10.0, 12.0, 9.0, dt=1e-3, fs=fs, nsponge=4, qMin=0.0, qInterior=0.0, nthread=4)
    for k, v in model.items():
        getattr(p, k)[...] = v
    return p


def _step(o, kind, fs, model, fields):
    p = _prop(o, kind, fs, model)
    for k, v in fields.items():
        getattr(p, k)[...] = v
    p.time_step()
    return p


def _fields(kind, fs, seed):
    rng = np.random.default_rng(seed)
    f = {"pCur": _interior(smooth(rng, SHAPE, 1.0, 2), fs), "pOld": _interior(smooth(rng, SHAPE, 1.0, 2), fs)}
    if kind != "iso":
        f["mCur"] = _interior(smooth(rng, SHAPE, 1.0, 2), fs)
        f["mOld"] = _interior(smooth(rng, SHAPE, 1.0, 2), fs)
    return f


def _fd(o, kind, fs, model, fields, dm, h):
    mp = {k: (v + (h * dm[k] if k in dm else 0)).astype(F32) for k, v in model.items()}
    mm = {k: (v - (h * dm[k] if k in dm else 0)).astype(F32) for k, v in model.items()}
    sp = _step(o, kind, fs, mp, fields)
    sm = _step(o, kind, fs, mm, fields)
    out = {"pCur": (sp.pCur.astype(np.float64) - sm.pCur) / (2 * h)}
    if kind != "iso":
        out["mCur"] = (sp.mCur.astype(np.float64) - sm.mCur) / (2 * h)
    return out


def _born(o, kind, fs, model, fields, dm, which):
    base = _step(o, kind, fs, model, fields)
    base.scale_spatial_derivatives()
    lin = _prop(o, kind, fs, model)
    P, DP = fields["pCur"], base.pSpace
    if kind == "iso":
        if which == "V":
            lin.forward_born_injection_V(dm["v"], DP)
        elif which == "VB":
            lin.forward_born_injection_VB(dm["v"], dm["b"], P, DP)
        else:
            lin.forward_born_injection_B(dm["b"], P, DP)
    else:
        M, DM = fields["mCur"], base.mSpace
        if which == "V":
            lin.forward_born_injection_V(dm["v"], DP, DM)
        else:
            lin.forward_born_injection_VEA(dm["v"], dm["eps"], dm["eta"], P, M, DP, DM)
    return lin


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [("iso", "V"), ("iso", "VB"), ("iso", "B"), ("vti", "V"), ("vti", "VEA"), ("tti", "V"), ("tti", "VEA")]


@pytest.mark.parametrize("kind,which", CASES)
@pytest.mark.parametrize("fs", [False, True])
def test_born_source_is_the_linearised_step(oracle, kind, which, fs):
    rng = np.random.default_rng(17)
    model = make_model(kind, SHAPE, 23)
    fields = _fields(kind, fs, 29)
    dm = {}
    if "V" in which:
        dm["v"] = smooth(rng, SHAPE, 100.0, 2)
    if "B" in which:
        dm["b"] = smooth(rng, SHAPE, 0.02, 2)
    if "EA" in which:
        dm["eps"] = smooth(rng, SHAPE, 0.02, 2)
        dm["eta"] = smooth(rng, SHAPE, 0.02, 2)
    # central differences at h and h/2 combined by Richardson extrapolation (O(h^4)); large
    # steps keep float32 cancellation small
    fd1 = _fd(oracle, kind, fs, model, fields, dm, 1.0)
    fd2 = _fd(oracle, kind, fs, model, fields, dm, 0.5)
    fd = {k: (4 * fd2[k] - fd1[k]) / 3 for k in fd1}
    lin = _born(oracle, kind, fs, model, fields, dm, which)
    assert _rel(lin.pCur.astype(np.float64), fd["pCur"]) < 1e-3
    if kind != "iso":
        assert _rel(lin.mCur.astype(np.float64), fd["mCur"]) < 1e-3


@pytest.mark.parametrize("kind,which", CASES)
@pytest.mark.parametrize("fs", [False, True])
def test_imaging_condition_is_adjoint_of_source(oracle, kind, which, fs):
    rng = np.random.default_rng(31)
    model = make_model(kind, SHAPE, 37)
    fields = _fields(kind, fs, 41)
    dt = F32(1e-3)
    base = _step(oracle, kind, fs, model, fields)
    base.scale_spatial_derivatives()
    P, DP = fields["pCur"], base.pSpace.copy()
    lam_p = _interior(smooth(rng, SHAPE, 1.0, 2), fs)
    lam_m = _interior(smooth(rng, SHAPE, 1.0, 2), fs)
    names = {"V": ["v"], "VB": ["v", "b"], "B": ["b"], "VEA": ["v", "eps", "eta"]}[which]
    dm = {k: _interior(smooth(rng, SHAPE, 1.0, 2), fs) for k in names}
    # forward: source with zero initial wavefield
    lin = _prop(oracle, kind, fs, model)
    img = _prop(oracle, kind, fs, model)
    img.pOld[...] = lam_p
    g = {k: np.zeros(SHAPE, F32) for k in names}
    if kind == "iso":
        if which == "V":
            lin.forward_born_injection_V(dm["v"], DP)
            img.adjoint_born_accumulation_V(g["v"], DP)
        elif which == "VB":
            lin.forward_born_injection_VB(dm["v"], dm["b"], P, DP)
            img.adjoint_born_accumulation_VB(g["v"], g["b"], P, DP)
        else:
            lin.forward_born_injection_B(dm["b"], P, DP)
            img.adjoint_born_accumulation_B(g["b"], P, DP)
        lhs = np.sum(lin.pCur.astype(np.float64) * lam_p * (model["b"] / model["v"].astype(np.float64) ** 2))
    else:
        M, DM = fields["mCur"], base.mSpace.copy()
        img.mOld[...] = lam_m
        if which == "V":
            lin.forward_born_injection_V(dm["v"], DP, DM)
            img.adjoint_born_accumulation_V(g["v"], DP, DM)
        else:
            lin.forward_born_injection_VEA(dm["v"], dm["eps"], dm["eta"], P, M, DP, DM)
            img.adjoint_born_accumulation_VEA(g["v"], g["eps"], g["eta"], P, M, DP, DM)
        w_ = model["b"] / model["v"].astype(np.float64) ** 2
        lhs = np.sum((lin.pCur.astype(np.float64) * lam_p + lin.mCur.astype(np.float64) * lam_m) * w_)
    rhs = float(dt) ** 2 * sum(np.sum(dm[k].astype(np.float64) * g[k]) for k in names)
    assert abs(lhs - rhs) <= 1e-4 * (abs(lhs) + abs(rhs)), (lhs, rhs)
