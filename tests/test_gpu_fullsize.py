"""GPU: parity at the workloads' full sizes -- Iso 256^3 (40-cell sponge), VTI 512^3 (40) and
TTI 800^3 (50), free surface, the launch configuration the step chooses by itself at that size
(TTI 800^3: rounds in step over 200-plane chunks) -- two steps from a seeded state, every output
bit-identical to the CPU oracle run on the host's threads.

The fields are built in one pass each, so that 800^3 stays within seconds: a field is the outer
product of an (x, y) map (smooth, asymmetric, with a random part) and a z profile, i.e.
f(x, y, z) = a(x, y) c(z); a transposed or shifted index shows as a mismatch."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32
CASES = [("iso", 256, 40, 12.5), ("vti", 512, 40, 12.5), ("tti", 800, 50, 12.5)]


def _xy(rng, n, lo, hi):
    x = np.arange(n) / n
    k1, k2, p1, p2 = rng.uniform(1, 3), rng.uniform(0.5, 2.5), rng.uniform(0, 6), rng.uniform(0, 6)
    a = 0.5 + 0.3 * np.sin(2 * np.pi * k1 * x + p1)[:, None] * np.cos(2 * np.pi * k2 * x + p2)[None, :]
    a = a + 0.1 * rng.random((n, n))  # in [0.2, 0.9]
    return (lo + (hi - lo) * a).astype(F32)


def _z(rng, n, lo, hi, noise=0.0):
    z = np.arange(n) / n
    c = 0.5 + 0.4 * np.sin(2 * np.pi * rng.uniform(0.5, 3) * z + rng.uniform(0, 6)) + noise * rng.standard_normal(n)
    return (lo + (hi - lo) * c).astype(F32)


def _fill(out, a, c):
    np.multiply(a[:, :, None], c[None, None, :], out=out)


def _build(orc, kind, n, rng):
    """model and state of the oracle object, in place (oracle layout (nx, ny, nz))"""
    _fill(orc.v, _xy(rng, n, 1.0, 1.5), np.linspace(1500, 3000, n, dtype=F32))  # 1500 .. 4500 m/s
    _fill(orc.b, _xy(rng, n, 0.8, 1.2), _z(rng, n, 0.45, 0.85))
    if kind != "iso":
        _fill(orc.eps, _xy(rng, n, 0.5, 1.0), _z(rng, n, 0.0, 0.3))
        _fill(orc.eta, _xy(rng, n, 0.5, 1.0), _z(rng, n, 0.0, 0.15))
        _fill(orc.f, _xy(rng, n, 0.9, 1.0), np.full(n, 0.85, F32))
    if kind == "tti":
        ang = np.empty_like(orc.v)
        _fill(ang, _xy(rng, n, -1.0, 1.0), _z(rng, n, -0.6, 0.6))  # tilt within +-45 degrees
        np.sin(ang, out=orc.sinTheta)
        np.cos(ang, out=orc.cosTheta)
        _fill(ang, _xy(rng, n, -1.0, 1.0), _z(rng, n, -3.1, 3.1))
        np.sin(ang, out=orc.sinPhi)
        np.cos(ang, out=orc.cosPhi)
        del ang
    for name in ("pCur", "pOld") + (("mCur", "mOld") if kind != "iso" else ()):
        _fill(getattr(orc, name), _xy(rng, n, -1.0, 1.0), _z(rng, n, -1.0, 1.0, noise=0.2))


@pytest.mark.timeout(400)  # 800^3: about a minute of field building and oracle steps on the host
@pytest.mark.parametrize("kind,n,nsponge,d", CASES)
def test_full_size_steps_bit_exact(oracle, kind, n, nsponge, d):
    import wavefd_amd as w
    from wavefd_amd import ginsu
    t0 = time.time()
    ocls = {"iso": oracle.OracleIso, "vti": oracle.OracleVTI, "tti": oracle.OracleTTI}[kind]
    orc = ocls(n, n, n, d, d, d, dt=1e-3, fs=True, nsponge=nsponge, freqQ=10.0, qMin=0.1, qInterior=100.0)
    _build(orc, kind, n, np.random.default_rng(n))
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[kind]
    prop = cls(nz=n, ny=n, nx=n, dz=d, dy=d, dx=d, dt=1e-3, nsponge=nsponge, freesurface=True, freqQ=10.0,
               qMin=0.1, qInterior=100.0)
    w.coherence(prop, "device")
    stream = prop._lib.GetStream(prop.p)
    fields = ("pCur", "pOld", "pSpace") + (("mCur", "mOld", "mSpace") if kind != "iso" else ())
    for name in orc.MODEL + fields[:2] + fields[3:5]:
        arr = getattr(orc, name)
        ginsu.copy(w.device_ptr(prop, name), arr.ctypes.data, arr.nbytes, stream)
    ginsu.sync(stream)
    cfg = w.fused_config(prop)
    print(f"{kind} {n}^3: built and uploaded in {time.time() - t0:.1f} s, launch {cfg}", flush=True)
    for _ in range(2):
        w.propagateforward(prop)
    w.synchronize(prop)
    t1 = time.time()
    for _ in range(2):
        orc.time_step()
    print(f"{kind} {n}^3: oracle 2 steps in {time.time() - t1:.1f} s", flush=True)
    if kind == "tti":
        assert cfg["lockstep"]  # the default configuration of this workload
    got = np.empty_like(orc.v)
    for name in fields:
        ginsu.copy(got.ctypes.data, w.device_ptr(prop, name), got.nbytes, stream)
        ginsu.sync(stream)
        ref = getattr(orc, name)
        assert np.array_equal(got, ref), name
        assert np.abs(ref).max() > 0, name
    w.free(prop)
