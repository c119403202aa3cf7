"""GPU: device-side injection/extraction and the time loops of the reference's tests.

* libspacetime's entry points on device fields equal the host path bit for bit;
* the registered-source fast path (TimeStepInjectExtract) equals the reference's loop
  (step, swap handles, injectdata!, extractdata!) run by the oracle;
* the reference's 3D loop of test/spacetime.jl:459-502, replayed through the coherent
  host API exactly as the Julia test writes it, equals the oracle bit for bit;
* the 3D off-grid vs on-grid accuracy test of test/spacetime.jl:459-502 (max rel diff < .1).
"""
import numpy as np
import pytest

from tests.helpers import host_view, make_fields, make_model, to_oracle

pytestmark = pytest.mark.gpu

F32 = np.float32


def test_device_inject_extract_equals_host(oracle):
    import wavefd_amd as w
    from tests.devmem import DeviceArray
    rng = np.random.default_rng(1)
    nz, ny, nx, nr, nt = 30, 31, 32, 9, 40
    pts = np.zeros(400, w.SourcePoint32)
    pts["iu"] = rng.integers(1, nz * ny * nx + 1, 400)
    pts["ir"] = rng.integers(1, nr + 1, 400)
    pts["c"] = rng.standard_normal(400).astype(F32)
    field = rng.standard_normal(nz * ny * nx).astype(F32)
    data = rng.standard_normal((nr, nt)).astype(F32)
    for it in (1, 13, 40):
        a = field.copy()
        w.injectdata(a, w.source_blocking(pts, 4), data, it)
        dev = DeviceArray(field)
        w.injectdata(dev, w.source_blocking(pts, 4), data, it)
        assert np.array_equal(dev.numpy(), a)
        d1 = data.copy()
        w.extractdata(d1, field, it, w.receiver_blocking(pts, 4))
        d2 = DeviceArray(data)
        w.extractdata(d2, DeviceArray(field), it, w.receiver_blocking(pts, 4))
        assert np.array_equal(d2.numpy(), d1)
        d3 = data.copy()  # device field, host data: the call's time sample staged
        w.extractdata(d3, DeviceArray(field), it, w.receiver_blocking(pts, 4))
        assert np.array_equal(d3, d1)


def test_device_calls_follow_changed_points_and_reject_bad_time_index():
    # a later call with the same number of points but other coefficients or indices must not
    # reuse the earlier call's grouped points; a time index outside 1..nt changes nothing
    import wavefd_amd as w
    from tests.devmem import DeviceArray
    from wavefd_amd import _lib
    rng = np.random.default_rng(4)
    n, nr, nt = 2000, 6, 9
    pts = np.zeros(50, w.SourcePoint32)
    pts["iu"] = rng.integers(1, n + 1, 50)
    pts["ir"] = rng.integers(1, nr + 1, 50)
    pts["c"] = rng.standard_normal(50).astype(F32)
    field = rng.standard_normal(n).astype(F32)
    data = rng.standard_normal((nr, nt)).astype(F32)
    for change in ("none", "c", "iu", "ir"):
        q = pts.copy()
        if change == "c":
            q["c"][7] += 1.0
        elif change == "iu":
            q["iu"][3] = q["iu"][3] % n + 1
        elif change == "ir":
            q["ir"][5] = q["ir"][5] % nr + 1
        a = field.copy()
        w.injectdata(a, w.source_blocking(q, 3), data, 4)
        dev = DeviceArray(field)
        w.injectdata(dev, w.source_blocking(q, 3), data, 4)
        assert np.array_equal(dev.numpy(), a), change
        d1 = data.copy()
        w.extractdata(d1, field, 4, w.receiver_blocking(q, 3))
        d2 = data.copy()
        w.extractdata(d2, DeviceArray(field), 4, w.receiver_blocking(q, 3))
        assert np.array_equal(d2, d1), change
    dev = DeviceArray(field)
    for it in (0, nt + 1):
        w.injectdata(dev, w.source_blocking(pts, 3), DeviceArray(data), it)
        assert "time index" in _lib.last_error("spacetime")
    assert np.array_equal(dev.numpy(), field)


def test_device_inject_rejects_points_outside_field():
    import wavefd_amd as w
    from tests.devmem import DeviceArray
    from wavefd_amd import _lib
    pts = np.zeros(1, w.SourcePoint32)
    pts["iu"] = 1001
    pts["ir"] = 1
    pts["c"] = 1
    dev = DeviceArray(np.zeros(1000, F32))
    w.injectdata(dev, [pts], np.ones((1, 5), F32), 1)
    assert "outside" in _lib.last_error("spacetime")
    assert float(np.abs(dev.numpy()).sum()) == 0.0


@pytest.mark.parametrize("kind", ["iso", "vti", "tti"])
def test_registered_sources_receivers_match_reference_loop(oracle, kind):
    import wavefd_amd as w
    shape = (40, 38, 44)
    nx, ny, nz = shape
    d = (10.0, 10.0, 10.0)
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[kind]
    ocls = {"iso": oracle.OracleIso, "vti": oracle.OracleVTI, "tti": oracle.OracleTTI}[kind]
    prop = cls(nz=nz, ny=ny, nx=nx, dz=10.0, dy=10.0, dx=10.0, dt=1e-3, nsponge=8, freesurface=True, freqQ=10.0,
               qMin=0.5, qInterior=80.0)
    orc = ocls(nx, ny, nz, *d, dt=1e-3, fs=True, nsponge=8, freqQ=10.0, qMin=0.5, qInterior=80.0, nthread=8)
    for name, arr in make_model(kind, shape, 3).items():
        host_view(prop, name)[:] = arr.ravel()
        getattr(orc, name)[...] = arr
    nt = 60
    src = w.hickscoeffs(F32(10), F32(10), F32(10), 0.0, 0.0, 0.0, nz, ny, nx, [103.0], [197.5], [203.3])
    wavelet = w.causal_ricker(25.0, 1e-3 * np.arange(nt))[None, :]
    zr = np.full(30, 50.0)
    yr = np.linspace(60, 310, 30)
    xr = np.linspace(70, 320, 30)
    rcv = w.linearcoeffs(F32(10), F32(10), F32(10), 0.0, 0.0, 0.0, nz, ny, nx, zr, yr, xr)
    w.coherence(prop, "device")
    w.sync_to_device(prop)
    w.set_sources(prop, w.source_blocking(src, 4), wavelet, fields=3 if kind != "iso" else 1)
    w.set_receivers(prop, w.receiver_blocking(rcv, 4), nt)
    traces = np.zeros((30, nt), F32)
    for it in range(1, nt + 1):
        w.timestep_inject_extract(prop, it)
        orc.time_step()
        oracle.inject(orc.pCur, wavelet, it, np.concatenate(w.source_blocking(src, 4)))
        if kind != "iso":
            oracle.inject(orc.mCur, wavelet, it, np.concatenate(w.source_blocking(src, 4)))
        oracle.extract(traces, orc.pCur, it, np.concatenate(w.receiver_blocking(rcv, 4)))
    got = w.receiver_data(prop, 30, nt)
    w.sync_to_host(prop)
    assert np.array_equal(to_oracle(prop, "pCur", shape), orc.pCur)
    assert np.array_equal(got, traces)
    assert np.abs(traces).max() > 0


def test_julia_loop_replay_coherent(oracle):
    """test/spacetime.jl:470-492 written against the Python mirror of the Julia API."""
    import wavefd_amd as w
    nz = ny = nx = 41
    dz = dy = dx = 0.02
    dt, nt, nsponge = 0.001, 80, 10
    prop = w.Prop3DAcoIsoDenQ_DEO2_FDTD(nz=nz, ny=ny, nx=nx, nsponge=nsponge, dz=dz, dy=dy, dx=dx, dt=dt,
                                         freesurface=True)
    w.V(prop)[...] = 1.5
    w.B(prop)[...] = 1.0
    pcur = w.PCur(prop)
    pold = w.POld(prop)
    fpeak = 1.5 / (5 * dx) / 3
    wavelet = w.causal_ricker(fpeak, dt * np.arange(nt))[None, :].astype(F32)
    points = w.hickscoeffs(F32(dz), F32(dy), F32(dx), 0.0, 0.0, 0.0, nz, ny, nx, [0.41], [0.41], [0.41])
    blocks = w.source_blocking(points, 8)
    orc = oracle.OracleIso(nx, ny, nz, dx, dy, dz, dt=dt, fs=True, nsponge=nsponge, freqQ=5.0, qMin=0.1,
                           qInterior=100.0, nthread=8)
    orc.v[...] = 1.5
    orc.b[...] = 1.0
    for it in range(1, nt + 1):
        w.propagateforward(prop)
        pcur, pold = pold, pcur
        w.injectdata(pcur, blocks, wavelet, it)
        orc.time_step()
        oracle.inject(orc.pCur, wavelet, it, np.concatenate(blocks))
    got = np.ascontiguousarray(pcur.transpose(2, 1, 0))
    assert np.array_equal(got, orc.pCur)
    assert np.abs(got).max() > 0


@pytest.mark.parametrize("F", ["hicks", "linear"])
@pytest.mark.parametrize("fs", [True, False])
def test_offgrid_vs_ongrid_accuracy_3d(F, fs):
    """test/spacetime.jl:459-502: a half-resolution off-grid run vs an on-grid run."""
    import wavefd_amd as w

    def modeling(ongrid):
        z = y = x = 2.0
        tmax, dt, nsponge = 0.5, 0.001, 10
        d = 0.01 if ongrid else 0.02
        fpeak = 1.5 / (5 * 2 * d) / 3 if ongrid else 1.5 / (5 * d) / 3
        n = int(round(z / d)) + 1
        nt = int(round(tmax / dt)) + 1
        s = z / 2 + 0.01
        prop = w.Prop3DAcoIsoDenQ_DEO2_FDTD(nz=n, ny=n, nx=n, nsponge=nsponge, dz=d, dy=d, dx=d, dt=dt,
                                             freesurface=fs)
        w.V(prop)[...] = 1.5
        w.B(prop)[...] = 1.0
        wavelet = w.causal_ricker(fpeak, dt * np.arange(nt))[None, :].astype(F32)
        fn = w.hickscoeffs if F == "hicks" else w.linearcoeffs
        points = fn(F32(d), F32(d), F32(d), 0.0, 0.0, 0.0, n, n, n, [s], [s], [s])
        w.coherence(prop, "device")
        w.sync_to_device(prop)
        w.set_sources(prop, w.source_blocking(points, 8), wavelet)
        for it in range(1, nt + 1):
            w.timestep_inject_extract(prop, it)
        w.sync_to_host(prop)
        p = np.array(w.PCur(prop)) / (d * d * d)
        w.free(prop)
        return p

    poff = modeling(False)
    pon = modeling(True)[::2, ::2, ::2]
    r = np.max((poff - pon) / np.abs(pon).max())
    assert r < 0.1


def test_device_inject_heavy_collisions_and_empty():
    # 600 taps on 16 cells (every cell shared by ~37 taps, many blockings): the per-cell sums
    # must run in the host path's order; an empty point list changes nothing
    import wavefd_amd as w
    from tests.devmem import DeviceArray
    rng = np.random.default_rng(7)
    n, nr, nt = 4096, 5, 12
    pts = np.zeros(600, w.SourcePoint32)
    cells = rng.choice(np.arange(1, n + 1), 16, replace=False)
    pts["iu"] = rng.choice(cells, 600)
    pts["ir"] = rng.integers(1, nr + 1, 600)
    pts["c"] = rng.standard_normal(600).astype(F32)
    field = rng.standard_normal(n).astype(F32)
    data = rng.standard_normal((nr, nt)).astype(F32)
    for nthreads in (1, 3, 8):
        for it in (1, 7, nt):
            a = field.copy()
            w.injectdata(a, w.source_blocking(pts, nthreads), data, it)
            dev = DeviceArray(field)
            w.injectdata(dev, w.source_blocking(pts, nthreads), data, it)
            assert np.array_equal(dev.numpy(), a), (nthreads, it)
    dev = DeviceArray(field)
    w.injectdata(dev, [], data, 1)
    assert np.array_equal(dev.numpy(), field)
    d2 = DeviceArray(data)
    w.extractdata(d2, DeviceArray(field), 1, [])
    assert np.array_equal(d2.numpy(), data)
