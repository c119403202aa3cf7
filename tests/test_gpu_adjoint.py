"""GPU: the adjoint dot-product test and the 2000-step parity of each physics.

Dot test.  One forward time step (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:247-261 and its VTI/TTI
twins) is u+ = (M L + 2 - Q) u - (1 - Q) u-, with M = dt^2 v^2 / b and L = D-(S D+) where the
material sandwich S is symmetric and D- = -D+^T on the update region.  M L is therefore
self-adjoint in the inner product weighted by W = M^-1, Q is diagonal, and the whole time loop
F (inject s after each step, extract receivers) has the adjoint
    F* d = S^T W  F_reversed( W^-1 R^T d reversed in time ),
i.e. the same propagator run on time-reversed receiver data injected with weights M at the
receivers and extracted with weights W at the sources (the reason the reference's
propagateadjoint! is the forward step, src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:69).  The test checks
<F s, d> = <s, F* d> to 1e-5 for seeded Gaussian s and d.

2000 steps.  A Hicks-interpolated Ricker source and trilinear receivers through the registered
fast path, 2000 steps, against the oracle running the reference's loop (step, inject,
extract): relative L2 of the wavefield and of the traces <= 1e-5 (they are in fact identical).
"""
import numpy as np
import pytest

from tests.helpers import host_view, make_model, rel_l2, to_oracle

pytestmark = pytest.mark.gpu

F32 = np.float32
CLS = {"iso": "Prop3DAcoIsoDenQ_DEO2_FDTD", "vti": "Prop3DAcoVTIDenQ_DEO2_FDTD", "tti": "Prop3DAcoTTIDenQ_DEO2_FDTD"}


def _points(rng, shape, n, margin=6):
    """n distinct on-grid interior points (never in the annulus nor on the surface row):
    SourcePoint32 with c = 1, 1-based Julia linear index iu and trace index ir."""
    import wavefd_amd as w
    nx, ny, nz = shape
    seen = set()
    pts = np.zeros(n, w.SourcePoint32)
    k = 0
    while k < n:
        ix, iy, iz = (int(rng.integers(margin, s - margin)) for s in (nx, ny, nz))
        if (ix, iy, iz) in seen:
            continue
        seen.add((ix, iy, iz))
        pts[k]["iu"] = iz + nz * (iy + ny * ix) + 1
        pts[k]["ir"] = k + 1
        pts[k]["c"] = 1.0
        k += 1
    return pts


def _cell(model, pts, shape):
    nx, ny, nz = shape
    i = pts["iu"] - 1
    ix, rem = np.divmod(i, ny * nz)
    iy, iz = np.divmod(rem, nz)
    return ix, iy, iz


def _run(kind, shape, model, dt, fs, src, src_data, rcv, nt):
    import wavefd_amd as w
    nx, ny, nz = shape
    prop = getattr(w, CLS[kind])(nz=nz, ny=ny, nx=nx, dz=10.0, dy=10.0, dx=10.0, dt=dt, nsponge=8,
                                 freesurface=fs, freqQ=10.0, qMin=0.5, qInterior=80.0)
    for name, arr in model.items():
        host_view(prop, name)[:] = arr.ravel()
    w.coherence(prop, "device")
    w.sync_to_device(prop)
    w.set_sources(prop, [src], src_data, fields=1)
    w.set_receivers(prop, [rcv], nt, fields=1)
    for it in range(1, nt + 1):
        w.timestep_inject_extract(prop, it)
    out = w.receiver_data(prop, len(rcv), nt).astype(np.float64)
    w.free(prop)
    return out


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
@pytest.mark.parametrize("fs", [True, False])
def test_adjoint_dot_product(kind, fs):
    shape = (36, 34, 40)
    nt, ns, nr = 300, 5, 7
    dt = 1e-3
    rng = np.random.default_rng({"iso": 11, "vti": 12, "tti": 13}[kind] + (100 if fs else 0))
    model = make_model(kind, shape, 5)
    src = _points(rng, shape, ns)
    rcv = _points(rng, shape, nr)
    s = rng.standard_normal((ns, nt)).astype(F32)
    d = rng.standard_normal((nr, nt)).astype(F32)
    # M = dt^2 v^2 / b at a point, W = 1 / M
    mw = {}
    for key, pts in (("src", src), ("rcv", rcv)):
        ix, iy, iz = _cell(model, pts, shape)
        v = model["v"][ix, iy, iz].astype(np.float64)
        b = model["b"][ix, iy, iz].astype(np.float64)
        mw[key] = dt * dt * v * v / b
    # forward
    fs_ = _run(kind, shape, model, dt, fs, src, s, rcv, nt)
    lhs = float(np.sum(fs_ * d))
    # adjoint: receivers become sources with weight M, sources become receivers with weight W
    asrc = rcv.copy()
    asrc["c"] = mw["rcv"].astype(F32)
    arcv = src.copy()
    arcv["c"] = (1.0 / mw["src"]).astype(F32)
    adj = _run(kind, shape, model, dt, fs, asrc, np.ascontiguousarray(d[:, ::-1]), arcv, nt)[:, ::-1]
    rhs = float(np.sum(s.astype(np.float64) * adj))
    assert abs(lhs) > 0
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs)), (lhs, rhs, abs(lhs - rhs) / abs(lhs))


# the second shape spans several tiles along z and y and several x-chunks of the default launch
# (3 z-tiles; 5 y-tiles of the two-wavefield wide tile, 9 of Iso's two-row tile)
@pytest.mark.parametrize("shape", [(30, 28, 34), (96, 80, 128)])
@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_2000_steps_match_oracle(oracle, kind, shape):
    import wavefd_amd as w
    nx, ny, nz = shape
    nt, dt, d = 2000, 1e-3, 10.0
    model = make_model(kind, shape, 21)
    prop = getattr(w, CLS[kind])(nz=nz, ny=ny, nx=nx, dz=d, dy=d, dx=d, dt=dt, nsponge=8, freesurface=True,
                                 freqQ=10.0, qMin=0.5, qInterior=80.0)
    ocls = {"iso": oracle.OracleIso, "vti": oracle.OracleVTI, "tti": oracle.OracleTTI}[kind]
    orc = ocls(nx, ny, nz, d, d, d, dt=dt, fs=True, nsponge=8, freqQ=10.0, qMin=0.5, qInterior=80.0, nthread=16)
    for name, arr in model.items():
        host_view(prop, name)[:] = arr.ravel()
        getattr(orc, name)[...] = arr
    src = w.hickscoeffs(F32(d), F32(d), F32(d), 0.0, 0.0, 0.0, nz, ny, nx, [52.0], [141.5 * ny / 28], [163.3 * nx / 30])
    wavelet = w.causal_ricker(15.0, dt * np.arange(nt))[None, :].astype(F32)
    nr = 24
    rcv = w.linearcoeffs(F32(d), F32(d), F32(d), 0.0, 0.0, 0.0, nz, ny, nx, np.full(nr, 25.0),
                         np.linspace(50, 220 * ny / 28, nr), np.linspace(60, 270 * nx / 30, nr))
    sblocks, rblocks = w.source_blocking(src, 4), w.receiver_blocking(rcv, 4)
    w.coherence(prop, "device")
    w.sync_to_device(prop)
    w.set_sources(prop, sblocks, wavelet, fields=3 if kind != "iso" else 1)
    w.set_receivers(prop, rblocks, nt)
    traces = np.zeros((nr, nt), F32)
    allsrc, allrcv = np.concatenate(sblocks), np.concatenate(rblocks)
    for it in range(1, nt + 1):
        w.timestep_inject_extract(prop, it)
        orc.time_step()
        oracle.inject(orc.pCur, wavelet, it, allsrc)
        if kind != "iso":
            oracle.inject(orc.mCur, wavelet, it, allsrc)
        oracle.extract(traces, orc.pCur, it, allrcv)
    got = w.receiver_data(prop, nr, nt)
    w.sync_to_host(prop)
    field = to_oracle(prop, "pCur", shape)
    assert np.abs(traces).max() > 0 and np.isfinite(traces).all()
    assert rel_l2(field, orc.pCur) <= 1e-5
    assert rel_l2(got, traces) <= 1e-5
    # the kernels evaluate the reference's expressions in its order: identical results
    assert np.array_equal(field, orc.pCur) and np.array_equal(got, traces)
    w.free(prop)
