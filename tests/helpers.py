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


def smooth_noise_big(shape, cells, seed):
    """Seeded smooth field in [-1, 1] on a large (nx, ny, nz) grid: U(-1, 1) noise on a coarse grid
    of one node per `cells` cells, interpolated trilinearly (three matrix products, float32), scaled
    to max |x| = 1.  A cheap stand-in for the Gaussian smoothing of the config models at sizes where
    filtering the full grid on the host would dominate a test."""
    rng = np.random.default_rng(seed)
    coarse_shape = [max(2, int(np.ceil(n / cells)) + 1) for n in shape]
    c = rng.uniform(-1, 1, coarse_shape).astype(F32)

    def interp(n, m):
        # (n, m) matrix of linear-interpolation weights from m coarse nodes to n points
        t = np.arange(n) * (m - 1) / max(n - 1, 1)
        i0 = np.minimum(np.floor(t).astype(int), m - 2)
        w1 = (t - i0).astype(F32)
        a = np.zeros((n, m), F32)
        a[np.arange(n), i0] = 1 - w1
        a[np.arange(n), i0 + 1] += w1
        return a

    ax, ay, az = (interp(n, m) for n, m in zip(shape, coarse_shape))
    x = np.tensordot(ax, c, axes=(1, 0))                      # (nx, my, mz)
    x = np.einsum("jb,ibc->ijc", ay, x, optimize=True)         # (nx, ny, mz)
    out = (x.reshape(-1, x.shape[2]) @ az.T).reshape(shape)   # (nx, ny, nz)
    out /= np.abs(out).max()
    return out.astype(F32)


def config_model(physics, shape, d, dt, seed):
    """The benchmark model construction in oracle layout (nx, ny, nz): v(z) rising linearly from
    1500 m/s at the top to 4500 m/s at the bottom times 1 +- 5 % smooth noise, eps in [0, 0.3] and
    eta in [0, 0.15] (smooth), f = 0.85, velocity capped so that v sqrt(1 + 2 eps) dt / d <= 0.42,
    density from Gardner's relation clamped to [1, 2.5] g/cc, b = 1 / rho."""
    nx, ny, nz = shape
    z = np.arange(nz, dtype=np.float64) * d
    vz = (1500.0 + 3000.0 * z / ((nz - 1) * d)).astype(F32)
    v = vz[None, None, :] * (1 + 0.05 * smooth_noise_big(shape, 15, seed))
    m = {}
    cap = F32(0.42 * d / dt)
    if physics != "iso":
        m["eps"] = (0.15 + 0.15 * smooth_noise_big(shape, 15, seed + 100)).astype(F32)
        m["eta"] = (0.075 + 0.075 * smooth_noise_big(shape, 15, seed + 200)).astype(F32)
        m["f"] = np.full(shape, 0.85, F32)
        v = np.minimum(v, cap / np.sqrt(1 + 2 * m["eps"]))
    else:
        v = np.minimum(v, cap)
    m["v"] = v.astype(F32)
    rho = np.clip(0.31 * m["v"] ** 0.25, 1.0, 2.5)
    m["b"] = (1.0 / rho).astype(F32)
    return m


def config_angles(shape, seed):
    """The TTI workload's angles in oracle layout: tilt theta in [-45, 45] deg smoothed by 25 cells,
    azimuth phi in [0, 360) deg as a unit vector (cos, sin) of two fields smoothed by 25 cells."""
    th = (np.pi / 4) * smooth_noise_big(shape, 25, seed)
    m = {"sinTheta": np.sin(th).astype(F32), "cosTheta": np.cos(th).astype(F32)}
    del th
    c = smooth_noise_big(shape, 25, seed + 1)
    s = smooth_noise_big(shape, 25, seed + 2)
    r = np.sqrt(c * c + s * s)
    np.maximum(r, 1e-6, out=r)
    m["cosPhi"] = (c / r).astype(F32)
    m["sinPhi"] = (s / r).astype(F32)
    return m


def config_q(q, nsponge, dt, seed, fs=True):
    """In place: Q from 30 to 200 in the interior (115 + 85 times noise smoothed by 15 cells, at
    10 Hz: dtOmegaInvQ = dt 2 pi 10 / Q) where the sponge distance is at least nsponge; q keeps
    SetupDtOmegaInvQ's profile elsewhere (distance as src/absorb.h:133-145)."""
    nx, ny, nz = q.shape
    dx_, dy_, kz = np.arange(nx), np.arange(ny), np.arange(nz)
    dx_ = np.minimum(dx_, nx - 1 - dx_)
    dy_ = np.minimum(dy_, ny - 1 - dy_)
    dz_ = nz - 1 - kz if fs else np.minimum(kz, nz - 1 - kz)
    ks = np.minimum(np.minimum(dx_[:, None, None], dy_[None, :, None]), dz_[None, None, :])
    qi = (dt * 2 * np.pi * 10.0 / (115.0 + 85.0 * smooth_noise_big(q.shape, 15, seed))).astype(F32)
    np.copyto(q, qi, where=ks >= nsponge)
