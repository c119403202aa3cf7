"""Shared test helpers: seeded models, and moving arrays between the oracle's C-ordered
(nx, ny, nz) layout and the library's Julia-ordered host views (same memory order)."""
import os

import numpy as np

F32 = np.float32

MODEL = {
    "iso": ("v", "b"),
    "vti": ("v", "eps", "eta", "b", "f"),
    "tti": ("v", "eps", "eta", "b", "f", "sinTheta", "cosTheta", "sinPhi", "cosPhi"),
}
GETTER = {"v": "V", "b": "B", "eps": "Eps", "eta": "Eta", "f": "F", "sinTheta": "SinTheta",
          "cosTheta": "CosTheta", "sinPhi": "SinPhi", "cosPhi": "CosPhi", "q": "DtOmegaInvQ",
          "pCur": "PCur", "pOld": "POld", "mCur": "MCur", "mOld": "MOld", "pSpace": "PSpace", "mSpace": "MSpace"}


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def smooth(rng, shape, amp, passes=3):
    x = rng.standard_normal(shape)
    for ax in range(len(shape)):
        for _ in range(passes):
            x = 0.25 * np.roll(x, 1, ax) + 0.5 * x + 0.25 * np.roll(x, -1, ax)
    return (amp * x / np.abs(x).max()).astype(F32)


def make_model(kind, shape, seed):
    """Seeded smooth model in oracle layout (nx, ny, nz), stable for dt=1ms, d~10m."""
    rng = np.random.default_rng(seed)
    m = {"v": (2000 + smooth(rng, shape, 500)).astype(F32),
         "b": (0.8 + smooth(rng, shape, 0.3)).astype(F32)}
    if kind != "iso":
        m["eps"] = (0.1 + smooth(rng, shape, 0.08)).astype(F32)
        m["eta"] = (0.05 + smooth(rng, shape, 0.04)).astype(F32)
        m["f"] = np.full(shape, 0.85, F32)
    if kind == "tti":
        th = smooth(rng, shape, 0.7)
        ph = smooth(rng, shape, 3.0)
        m["sinTheta"] = np.sin(th).astype(F32)
        m["cosTheta"] = np.cos(th).astype(F32)
        m["sinPhi"] = np.sin(ph).astype(F32)
        m["cosPhi"] = np.cos(ph).astype(F32)
    return m


def make_fields(kind, shape, seed):
    rng = np.random.default_rng(seed + 1000)
    f = {"pCur": smooth(rng, shape, 1.0, 2), "pOld": smooth(rng, shape, 1.0, 2)}
    if kind != "iso":
        f["mCur"] = smooth(rng, shape, 1.0, 2)
        f["mOld"] = smooth(rng, shape, 1.0, 2)
    return f


def host_view(prop, name):
    """Flat (memory-order) view of a library host array, equal to oracle C-order ravel."""
    import wavefd_amd as w
    arr = getattr(w, GETTER[name])(prop)
    return arr.ravel(order="F")


def to_oracle(prop, name, shape):
    return host_view(prop, name).reshape(shape).copy()


def rel_l2(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def ulp_diff(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    return int(np.abs(ai - bi).max()) if a.size else 0
