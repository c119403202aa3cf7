"""GPU: the HIP path at the benchmark workloads (parity by size-independent properties).

* Iso 256^3 at 10 m, 2000 steps with a Hicks point source and 256 x 256 receivers at 20 m depth:
  traces and wavefields bit-identical to the oracle running the same loop on the host;
* the TTI 800^3 workload's construction (tilt, azimuth, Q 30..200 as an array, 50-cell sponge) at
  200^3 over 2000 steps with a Hicks source into P and M and 200 x 200 receivers: bit-identical;
* VTI 512^3 at 12.5 m, 3000 steps of the linear step (Prop3DAcoVTIDenQ_DEO2_FDTD_TimeStepLinear,
  src/prop3DAcoVTIDenQ_DEO2_FDTD.h:356-376, the step the reference's propagateadjoint! runs): the
  adjoint dot-product test <F s, d> = <s, F* d> <= 1e-5 for seeded Gaussian source and receiver
  vectors, with the weights dt^2 v^2 / b and its inverse at the points (the construction of
  tests/test_gpu_adjoint.py), on the benchmark's seeded model;
* batched Born, Iso and VTI, on ginsu windows of a 1000 x 1000 x 600 model at 12.5 m (a 200^3
  window and a 600^3 one, nt = 40): two shots in one launch per step give the images and Born
  data of the two shots run one after the other, bit for bit; the Born dot test
  <J dv, d> = <dv, J^T d> <= 1e-5 on the 600^3 window (dv zero outside its interior box);
* VTI 512^3 Born gradient with the wave-predicted snapshot store (born.py::_WaveStore) against
  exact storage (the checkpointed store, which gives the exact store's image bit for bit, as
  tests/test_gpu_born_checkpoint.py checks): residuals quantised to 11 levels, >= 8x fewer stored
  bytes, and to 5 levels, >= 11x; image relative L2 <= 1e-3 for both.
  nt = 400 here (the store of 3000 frames would not fit one GPU beside the exact reference run).

The models follow the benchmark construction (tests/helpers.py::config_model).
"""
import numpy as np
import pytest

from tests.helpers import config_model, host_view, smooth_noise_big, to_oracle

pytestmark = pytest.mark.gpu

F32 = np.float32


def _on_grid_points(rng, shape, n, margin):
    import wavefd_amd as w
    nx, ny, nz = shape
    pts = np.zeros(n, w.SourcePoint32)
    seen = set()
    k = 0
    while k < n:
        c = tuple(int(rng.integers(margin, s - margin)) for s in (nx, ny, nz))
        if c in seen:
            continue
        seen.add(c)
        ix, iy, iz = c
        pts[k]["iu"] = iz + nz * (iy + ny * ix) + 1
        pts[k]["ir"] = k + 1
        pts[k]["c"] = 1.0
        k += 1
    return pts


def _weights(model, pts, shape, dt):
    nx, ny, nz = shape
    ix, rem = np.divmod(pts["iu"] - 1, ny * nz)
    iy, iz = np.divmod(rem, nz)
    v = model["v"][ix, iy, iz].astype(np.float64)
    b = model["b"][ix, iy, iz].astype(np.float64)
    return dt * dt * v * v / b


@pytest.mark.timeout(900)  # the oracle's 2000 steps of 256^3 on the host: one to three minutes
def test_c1_iso_256_forward_2000_steps_bit_exact(oracle):
    """Iso 256^3 at 10 m, dt 1 ms, 2000 steps, Q = 100 at 10 Hz, 40-cell sponge, free surface; a
    Hicks point source at 20 m depth in the middle of the grid with a 15 Hz Ricker wavelet, and a
    receiver at 20 m depth on every (x, y) node (256 x 256).  The GPU runs the loop of
    test/spacetime.jl:486-492 as TimeStepInjectExtractGraph; the oracle runs the same loop
    (step, inject, extract) on the host.  Traces and final wavefields are bit-identical."""
    import time
    import wavefd_amd as w
    n, d, dt, nt = 256, 10.0, 1e-3, 2000
    model = config_model("iso", (n, n, n), d, dt, 1)
    prop = w.Prop3DAcoIsoDenQ_DEO2_FDTD(nz=n, ny=n, nx=n, dz=d, dy=d, dx=d, dt=dt, nsponge=40,
                                        freesurface=True, freqQ=10.0, qMin=0.1, qInterior=100.0)
    orc = oracle.OracleIso(n, n, n, d, d, d, dt=dt, fs=True, nsponge=40, freqQ=10.0, qMin=0.1, qInterior=100.0,
                           nthread=16)
    for name, arr in model.items():
        host_view(prop, name)[:] = arr.ravel()
        getattr(orc, name)[...] = arr
    del model
    mid = (n - 1) * d / 2
    src = w.hickscoeffs(F32(d), F32(d), F32(d), 0.0, 0.0, 0.0, n, n, n, [20.0], [mid], [mid])
    wavelet = w.causal_ricker(15.0, dt * np.arange(nt))[None, :].astype(F32)
    ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    rcv = np.zeros(n * n, w.SourcePoint32)
    rcv["iu"] = 2 + n * (iy.ravel() + n * ix.ravel()) + 1  # iz = 2: 20 m
    rcv["ir"] = np.arange(n * n) + 1
    rcv["c"] = 1.0
    sparts, rparts = w.source_blocking(src, 16), w.receiver_blocking(rcv, 16)
    w.coherence(prop, "device")
    w.sync_to_device(prop)
    w.set_sources(prop, sparts, wavelet, fields=1)
    w.set_receivers(prop, rparts, nt)
    t0 = time.time()
    w.timestep_inject_extract_graph(prop, 1, nt)
    got = w.receiver_data(prop, n * n, nt)
    w.sync_to_host(prop)
    p_gpu = {name: to_oracle(prop, name, (n, n, n)).copy() for name in ("pCur", "pOld")}
    w.free(prop)
    print(f"C1 Iso 256^3: {nt} GPU steps with injection and extraction in {time.time() - t0:.1f} s", flush=True)
    spts, rpts = np.concatenate(sparts), np.concatenate(rparts)
    traces = np.zeros((n * n, nt), F32)
    t0 = time.time()
    for it in range(1, nt + 1):
        orc.time_step()
        oracle.inject(orc.pCur, wavelet, it, spts)
        oracle.extract(traces, orc.pCur, it, rpts)
        if it % 250 == 0:
            print(f"C1 oracle step {it} / {nt}, {time.time() - t0:.0f} s", flush=True)
    assert np.abs(traces).max() > 0 and np.isfinite(traces).all()
    assert np.array_equal(got, traces)
    assert np.array_equal(p_gpu["pCur"], orc.pCur)
    assert np.array_equal(p_gpu["pOld"], orc.pOld)


@pytest.mark.timeout(900)  # the oracle's 2000 TTI steps of 200^3 on the host: about a minute
def test_c3_construction_tti_200_forward_2000_steps_bit_exact(oracle):
    """The TTI workload's model construction at 200^3 (12.5 m, dt 1 ms, 50-cell sponge, free
    surface): VTI fields of the configs, tilt within +-45 deg and azimuth over 360 deg smoothed by
    25 cells, Q from 30 to 200 in the interior as an array (the sponge keeps SetupDtOmegaInvQ's
    profile).  A Hicks source at 25 m depth injected into P and M, a receiver at 25 m depth on
    every (x, y) node; 2000 steps of TimeStepInjectExtractGraph against the oracle's loop of
    test/spacetime.jl:486-492.  Traces and final wavefields are bit-identical."""
    import time
    import wavefd_amd as w
    from tests.helpers import config_angles, config_q
    n, d, dt, nt, nsponge = 200, 12.5, 1e-3, 2000, 50
    shape = (n, n, n)
    orc = oracle.OracleTTI(n, n, n, d, d, d, dt=dt, fs=True, nsponge=nsponge, freqQ=10.0, qMin=0.1,
                           qInterior=100.0, nthread=16)
    model = config_model("tti", shape, d, dt, 3)
    model.update(config_angles(shape, 303))
    config_q(orc.q, nsponge, dt, 306)
    model["q"] = orc.q.copy()
    prop = w.Prop3DAcoTTIDenQ_DEO2_FDTD(nz=n, ny=n, nx=n, dz=d, dy=d, dx=d, dt=dt, nsponge=nsponge,
                                        freesurface=True, freqQ=10.0, qMin=0.1, qInterior=100.0)
    for name, arr in model.items():
        host_view(prop, name)[:] = arr.ravel()
        getattr(orc, name)[...] = arr
    del model
    mid = (n - 1) * d / 2
    src = w.hickscoeffs(F32(d), F32(d), F32(d), 0.0, 0.0, 0.0, n, n, n, [25.0], [mid + 3.1], [mid - 4.7])
    wavelet = w.causal_ricker(15.0, dt * np.arange(nt))[None, :].astype(F32)
    ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    rcv = np.zeros(n * n, w.SourcePoint32)
    rcv["iu"] = 2 + n * (iy.ravel() + n * ix.ravel()) + 1  # iz = 2: 25 m
    rcv["ir"] = np.arange(n * n) + 1
    rcv["c"] = 1.0
    sparts, rparts = w.source_blocking(src, 16), w.receiver_blocking(rcv, 16)
    w.coherence(prop, "device")
    w.sync_to_device(prop)
    w.set_sources(prop, sparts, wavelet, fields=3)
    w.set_receivers(prop, rparts, nt)
    t0 = time.time()
    w.timestep_inject_extract_graph(prop, 1, nt)
    got = w.receiver_data(prop, n * n, nt)
    w.sync_to_host(prop)
    gpu = {name: to_oracle(prop, name, shape).copy() for name in ("pCur", "pOld", "mCur", "mOld")}
    w.free(prop)
    print(f"C3 construction, TTI 200^3: {nt} GPU steps with injection and extraction in {time.time() - t0:.1f} s",
          flush=True)
    spts, rpts = np.concatenate(sparts), np.concatenate(rparts)
    traces = np.zeros((n * n, nt), F32)
    t0 = time.time()
    for it in range(1, nt + 1):
        orc.time_step()
        oracle.inject(orc.pCur, wavelet, it, spts)
        oracle.inject(orc.mCur, wavelet, it, spts)
        oracle.extract(traces, orc.pCur, it, rpts)
        if it % 250 == 0:
            print(f"C3 construction oracle step {it} / {nt}, {time.time() - t0:.0f} s", flush=True)
    assert np.abs(traces).max() > 0 and np.isfinite(traces).all()
    assert np.array_equal(got, traces)
    for name, arr in gpu.items():
        assert np.array_equal(arr, getattr(orc, name)), name


def test_c2_vti_512_linear_step_dot_product():
    import wavefd_amd as w
    shape = (512, 512, 512)
    nx, ny, nz = shape
    d, dt, nt, ns, nr = 12.5, 1e-3, 3000, 6, 10
    model = config_model("vti", shape, d, dt, 2)
    rng = np.random.default_rng(2002)
    src = _on_grid_points(rng, shape, ns, 60)
    rcv = _on_grid_points(rng, shape, nr, 60)
    s = rng.standard_normal((ns, nt)).astype(F32)
    dd = rng.standard_normal((nr, nt)).astype(F32)
    prop = w.Prop3DAcoVTIDenQ_DEO2_FDTD(nz=nz, ny=ny, nx=nx, dz=d, dy=d, dx=d, dt=dt, nsponge=40,
                                        freesurface=True, freqQ=10.0, qMin=0.1, qInterior=100.0)
    for name, arr in model.items():
        host_view(prop, name)[:] = arr.ravel()
    w.coherence(prop, "device")
    w.sync_to_device(prop)
    zero = np.zeros((nx, ny, nz), F32)

    def run(spts, sdata, rpts):
        for name in ("pCur", "pOld", "mCur", "mOld"):
            host_view(prop, name)[:] = zero.ravel()
        w.sync_to_device(prop)
        w.set_sources(prop, [spts], sdata, fields=1)
        w.set_receivers(prop, [rpts], nt, fields=1)
        for it in range(1, nt + 1):
            w.propagateforward_linear(prop)
            assert prop._lib.InjectRegistered(prop.p, it) == 0
            assert prop._lib.ExtractRegistered(prop.p, it) == 0
        return w.receiver_data(prop, len(rpts), nt).astype(np.float64)

    fwd = run(src, s, rcv)
    lhs = float(np.sum(fwd * dd))
    asrc, arcv = rcv.copy(), src.copy()
    asrc["c"] = _weights(model, rcv, shape, dt).astype(F32)
    arcv["c"] = (1.0 / _weights(model, src, shape, dt)).astype(F32)
    adj = run(asrc, np.ascontiguousarray(dd[:, ::-1]), arcv)[:, ::-1]
    rhs = float(np.sum(s.astype(np.float64) * adj))
    w.free(prop)
    assert np.isfinite(fwd).all() and abs(lhs) > 0
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"C2 VTI 512^3, {nt} linear steps: <F s, d> = {lhs:.9e}, <s, F* d> = {rhs:.9e}, rel {rel:.2e}", flush=True)
    assert rel <= 1e-5


BIG = (600, 1000, 1000)  # (nz, ny, nx)


def _big_model(kind, d, dt):
    from wavefd_amd import born, ginsu
    nz, ny, nx = BIG
    m = config_model(kind, (nx, ny, nz), d, dt, 104)
    dev = {}
    for k in born.MODEL_FIELDS[kind]:
        dev[k] = ginsu.DeviceBuffer(m[k].nbytes)
        dev[k].upload(m[k])
    dv = (0.02 * m["v"] * smooth_noise_big((nx, ny, nz), 15, 204)).astype(F32)
    return m, dev, dv


def _shot(born, win, d, wav):
    oz, oy, ox = win.origin
    snz, sny, snx = win.size
    ry, rx = np.meshgrid(np.linspace((oy + 30) * d, (oy + sny - 30) * d, 16),
                         np.linspace((ox + 30) * d, (ox + snx - 30) * d, 16), indexing="ij")
    return born.BornShot((20.0, (oy + sny / 2.0) * d, (ox + snx / 2.0) * d), np.full(ry.size, 20.0), ry.ravel(),
                         rx.ravel(), wav)


@pytest.mark.parametrize("kind", ["iso", "vti"])
def test_c4_batched_born_on_200_and_600_cube_windows(kind):
    from wavefd_amd import born, ginsu
    d, dt, nt, nsponge = 12.5, 1e-3, 40, 30
    nz, ny, nx = BIG
    m, dev, dv = _big_model(kind, d, dt)
    windows = [ginsu.Window(BIG, (0, 100, 700), (200, 200, 200)), ginsu.Window(BIG, (0, 300, 50), (600, 600, 600))]
    wav = np.asarray(born.spacetime.causal_ricker(15.0, dt * np.arange(nt)), F32)
    shots = [_shot(born, wdw, d, wav) for wdw in windows]
    fields = {k: b.ptr for k, b in dev.items()}
    dvb = ginsu.DeviceBuffer(dv.nbytes)
    dvb.upload(dv)
    seq, bat = ginsu.DeviceBuffer(dv.nbytes), ginsu.DeviceBuffer(dv.nbytes)
    seq.zero()
    bat.zero()
    data_seq = [born.born_shot(kind, wdw, (d, d, d), dt, sh, fields, dvb.ptr, seq.ptr, nsponge=nsponge)["born_data"]
                for wdw, sh in zip(windows, shots)]
    data_bat = born.born_shots(kind, windows, (d, d, d), dt, shots, fields, dvb.ptr, bat.ptr, nsponge=nsponge,
                               batch=2)
    gs = seq.download(np.zeros((nx, ny, nz), F32))
    gb = bat.download(np.zeros((nx, ny, nz), F32))
    assert np.isfinite(gs).all() and np.abs(gs).max() > 0
    assert np.array_equal(gs, gb)
    for a, b in zip(data_seq, data_bat):
        assert np.abs(a).max() > 0 and np.array_equal(a, b)
    del gs, gb
    # dot test on the 600^3 window, dv restricted to the box its image covers
    w6 = windows[1]
    (lz0, lz1), (ly0, ly1), (lx0, lx1) = w6.interior(nsponge)
    oz, oy, ox = w6.origin
    dvm = np.zeros_like(dv)
    sl = (slice(ox + lx0, ox + lx1), slice(oy + ly0, oy + ly1), slice(oz + lz0, oz + lz1))
    dvm[sl] = dv[sl]
    dvb.upload(dvm)
    img = seq
    img.zero()
    rng = np.random.default_rng(44)
    d2 = rng.standard_normal((shots[1].nr, nt)).astype(F32)
    out = born.born_shot(kind, w6, (d, d, d), dt, shots[1], fields, dvb.ptr, img.ptr, nsponge=nsponge,
                         adjoint_data=d2)
    g = img.download(np.zeros((nx, ny, nz), F32))
    lhs = float(np.sum(out["born_data"].astype(np.float64) * d2))
    rhs = float(np.sum(dvm[sl].astype(np.float64) * g[sl]))
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"C4 {kind} 600^3 window, nt {nt}: <J dv, d> = {lhs:.9e}, <dv, J^T d> = {rhs:.9e}, rel {rel:.2e}",
          flush=True)
    assert abs(lhs) > 0 and rel <= 1e-5
    for b in list(dev.values()) + [dvb, seq, bat]:
        b.free()


def test_c5_vti_512_predictive_store_against_exact():
    from wavefd_amd import born, ginsu
    shape = (512, 512, 512)
    nx, ny, nz = shape
    d, dt, nt = 12.5, 1e-3, 400
    m = config_model("vti", shape, d, dt, 5)
    dev = {}
    for k in born.MODEL_FIELDS["vti"]:
        dev[k] = ginsu.DeviceBuffer(m[k].nbytes)
        dev[k].upload(m[k])
    dv = (0.02 * m["v"] * smooth_noise_big(shape, 15, 305)).astype(F32)
    dvb = ginsu.DeviceBuffer(dv.nbytes)
    dvb.upload(dv)
    win = ginsu.Window((nz, ny, nx), (0, 0, 0), (nz, ny, nx))
    wav = np.asarray(born.spacetime.causal_ricker(15.0, dt * np.arange(nt)), F32)
    shot = _shot(born, win, d, wav)
    fields = {k: b.ptr for k, b in dev.items()}
    imgs, stored = {}, {}
    for mode in ("exact", 11, 5):
        img = ginsu.DeviceBuffer(dv.nbytes)
        img.zero()
        r = born.born_shot("vti", win, (d, d, d), dt, shot, fields, dvb.ptr, img.ptr, nsponge=40,
                           keep_born_data=False, checkpoint=20 if mode == "exact" else None,
                           predictive=None if mode == "exact" else (96, "wave", 16, mode))
        if mode != "exact":
            stored[mode] = r["stored_bytes"]
        imgs[mode] = img.download(np.zeros(shape, F32)).astype(np.float64)
        img.free()
    for b in list(dev.values()) + [dvb]:
        b.free()
    assert np.isfinite(imgs["exact"]).all() and np.abs(imgs["exact"]).max() > 0
    for levels, want_ratio in ((11, 8.0), (5, 11.0)):
        ratio = 4.0 * 2 * nx * ny * nz * nt / stored[levels]
        rel = float(np.linalg.norm(imgs[levels] - imgs["exact"]) / np.linalg.norm(imgs["exact"]))
        print(f"C5 VTI 512^3, nt {nt}: wave-predicted store, residuals to {levels} levels: {ratio:.2f}x fewer bytes "
              f"than float32, image rel L2 {rel:.2e}", flush=True)
        assert ratio >= want_ratio
        assert rel <= 1e-3
