"""Throughput of the WaveFD 3D acoustic forward step on one MI355X.

Headline workload: 3D AcoTTIDenQ_DEO2_FDTD forward, FP32, 800^3 at 12.5 m, dt = 1 ms, 50-cell
sponge, free surface; VTI fields (seeded velocity ramp 1500->4500 m/s +-5% smoothed by a
15-cell Gaussian, Gardner density, eps in U(0,0.3), eta in U(0,0.15)), tilt in U(-45,45) deg
and azimuth in U(0,360) deg smoothed by a 25-cell Gaussian, Q varying 30..200 at 10 Hz.
Velocity is capped so that v*sqrt(1+2 eps)*dt/d <= 0.42 (the 8th-order staggered scheme's
3D limit is 0.4488).  One "step" is one TimeStep of the full grid, inputs resident in HBM.

Mcells/s = nx*ny*nz*steps / t / 1e6, the reference's definition (benchmark/mcells_per_second.jl:16).
Roofline: algorithmic HBM bytes of the fused step (TTI: 14 float reads + 4 float writes =
72 B/cell) / measured kernel time, against MI355X's 8 TB/s.

The default line also carries the Iso 256^3 and VTI 512^3 steps ("extra"), Born shots per hour of
the fused path against the reference's separate calls on 24 ginsu windows ("born"), and the Born
image with compressed snapshots against exact storage ("born.snapshot_compression").

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--physics tti|vti|iso] [--n N]
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Iso and VTI keep SetupDtOmegaInvQ's profile here and read Q from its table by sponge distance
# (no Q stream: Iso 6 float streams, VTI 13); TTI's workload writes its own Q (18 streams)
BYTES_PER_CELL = {"iso": 24, "vti": 52, "tti": 72}
HBM_PEAK_GBS = 8000.0
REFERENCE_44T_MCELLS = {"iso": 1961.90, "vti": 893.94, "tti": 800.58}  # docs/src/index.md:42-44


def _hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipMemcpy.restype = ctypes.c_int
    return lib


def _copy_d2d(hip, dst, src_tensor):
    rc = hip.hipMemcpy(dst, src_tensor.data_ptr(), src_tensor.numel() * 4, 3)
    if rc != 0:
        raise RuntimeError("hipMemcpy failed: %d" % rc)


def _smooth_noise(torch, shape, sigma, seed, dev):
    """Seeded U(-1,1) noise smoothed by three passes of a box filter per axis (a Gaussian of
    std ~sigma), reflecting edges, scaled to max |x| = 1.  Cumulative sums only."""
    g = torch.Generator(device=dev).manual_seed(seed)
    x = torch.rand(shape, generator=g, device=dev, dtype=torch.float32) * 2 - 1
    wdt = max(1, int(round((12.0 * sigma * sigma / 3.0 + 1.0) ** 0.5)))
    h = wdt // 2
    for axis in range(3):
        for _ in range(3):
            xt = x.movedim(axis, 0)
            n = xt.shape[0]
            idx = torch.arange(-h - 1, n + h, device=dev)
            idx = torch.where(idx < 0, -idx, idx)
            idx = torch.where(idx > n - 1, 2 * (n - 1) - idx, idx).clamp(0, n - 1)
            c = torch.cumsum(xt.index_select(0, idx).double(), 0)
            y = ((c[wdt:wdt + n] - c[0:n]) / wdt).float()
            x = y.movedim(0, axis).contiguous()
            del xt, c, y
    return x / x.abs().max()


def build_model(torch, physics, n, d, dt, seed, dev):
    """Return dict of (nx,ny,nz) float32 device tensors for the configured synthetic model
    (n: the cube's edge, or (nx, ny, nz))."""
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    shape = (nx, ny, nz)
    z = torch.arange(nz, device=dev, dtype=torch.float32) * d
    vz = 1500.0 + 3000.0 * z / ((nz - 1) * d)
    v = vz.view(1, 1, nz).expand(shape) * (1 + 0.05 * _smooth_noise(torch, shape, 15, seed, dev))
    m = {}
    if physics != "iso":
        m["eps"] = 0.15 + 0.15 * _smooth_noise(torch, shape, 15, seed + 100, dev)
        m["eta"] = 0.075 + 0.075 * _smooth_noise(torch, shape, 15, seed + 200, dev)
        m["f"] = torch.full(shape, 0.85, device=dev)
        vcap = 0.42 * d / (dt * torch.sqrt(1 + 2 * m["eps"]))
    else:
        vcap = torch.full(shape, 0.42 * d / dt, device=dev)
    v = torch.minimum(v, vcap).contiguous()
    rho = torch.clamp(0.31 * v ** 0.25, 1.0, 2.5)
    m["v"] = v
    m["b"] = (1.0 / rho).contiguous()
    if physics == "tti":
        th = (np.pi / 4) * _smooth_noise(torch, shape, 25, seed + 300, dev)
        c = _smooth_noise(torch, shape, 25, seed + 400, dev)
        s = _smooth_noise(torch, shape, 25, seed + 500, dev)
        nrm = torch.sqrt(c * c + s * s).clamp_min(1e-6)
        m["sinTheta"], m["cosTheta"] = torch.sin(th), torch.cos(th)
        m["sinPhi"], m["cosPhi"] = (s / nrm).contiguous(), (c / nrm).contiguous()
    return m


def _time_taps(torch, w, prop, physics, n, d, steps, warmup, stream):
    """The same handle's modelling loop: a Hicks source half a cell off the grid at 20 m depth
    in the middle, n x n receivers on the nodes of the row at 2d depth (one unit tap each), and
    per step TimeStepInjectExtract (the step, then the registered injection and extraction
    kernels on its stream).  HIP-event and wall time per step."""
    f32 = np.float32
    nt = warmup + steps
    mid = (n / 2 - 0.5) * d
    src = w.hickscoeffs(f32(d), f32(d), f32(d), 0.0, 0.0, 0.0, n, n, n, [20.0], [mid], [mid])
    y, x = np.meshgrid(np.arange(n) * d, np.arange(n) * d, indexing="ij")
    rcv = w.linearcoeffs(f32(d), f32(d), f32(d), 0.0, 0.0, 0.0, n, n, n, np.full(n * n, 2 * d), y.ravel(),
                         x.ravel())
    wavelet = w.causal_ricker(15.0, 1e-3 * np.arange(nt))[None, :]
    w.set_sources(prop, w.source_blocking(src, 8), wavelet, fields=1 if physics == "iso" else 3)
    w.set_receivers(prop, w.receiver_blocking(rcv, 8), nt)
    for it in range(1, warmup + 1):
        w.timestep_inject_extract(prop, it)
    w.synchronize(prop)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    ev0.record(stream)
    for it in range(warmup + 1, nt + 1):
        w.timestep_inject_extract(prop, it)
    ev1.record(stream)
    w.synchronize(prop)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    return {"ms_per_step": wall * 1e3 / steps, "gpu_ms_per_step": ev0.elapsed_time(ev1) / steps,
            "receivers": n * n, "receiver_taps": int(len(rcv)), "source_taps": int(len(src))}


def run_gpu(physics, n, steps, warmup, seed=3, tune=False, with_taps=False):
    import torch
    import wavefd_amd as w

    dev = torch.device("cuda:0")
    d, dt = 12.5, 1e-3
    nsponge = 50 if physics == "tti" else 40
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[physics]
    t0 = time.time()
    prop = cls(nz=n, ny=n, nx=n, dz=d, dy=d, dx=d, dt=dt, nsponge=nsponge, freesurface=True, freqQ=10.0,
               qMin=0.1, qInterior=100.0)
    w.coherence(prop, "device")
    hip = _hip()
    model = build_model(torch, physics, n, d, dt, seed, dev)
    for name, t in model.items():
        _copy_d2d(hip, w.device_ptr(prop, name), t)
    if physics == "tti":
        # Q varying 30..200 in the interior; the sponge keeps SetupDtOmegaInvQ's profile
        q = torch.empty((n, n, n), device=dev)
        hip.hipMemcpy(q.data_ptr(), w.device_ptr(prop, "dtOmegaInvQ"), q.numel() * 4, 3)
        qint = 115.0 + 85.0 * _smooth_noise(torch, (n, n, n), 15, seed + 600, dev)
        k = torch.arange(n, device=dev)
        dist = torch.minimum(k, n - 1 - k)
        ks = torch.minimum(torch.minimum(dist.view(-1, 1, 1), dist.view(1, -1, 1)), (n - 1 - k).view(1, 1, -1))
        q = torch.where(ks >= nsponge, dt * 2 * np.pi * 10.0 / qint, q).contiguous()
        _copy_d2d(hip, w.device_ptr(prop, "dtOmegaInvQ"), q)
        del q, qint
    # a smooth initial wavefield so the arithmetic is not on zeros
    p0 = _smooth_noise(torch, (n, n, n), 4, seed + 700, dev)
    for name in ("pCur", "pOld") + (("mCur", "mOld") if physics != "iso" else ()):
        _copy_d2d(hip, w.device_ptr(prop, name), p0)
    del model, p0
    torch.cuda.synchronize()
    setup_s = time.time() - t0

    tuned = None
    if tune:
        # the launch-configuration autotuner (setup, before the timed region; it restores the
        # wavefields it steps)
        best, table = w.autotune(prop, nsteps=4, warmup=2)
        tuned = {"best": best, "candidates": [list(r) for r in table]}
    launch = w.fused_config(prop)
    stream = torch.cuda.ExternalStream(prop._lib.GetStream(prop.p))
    for _ in range(warmup):
        w.propagateforward(prop)
    w.synchronize(prop)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        w.propagateforward(prop)
    ev1.record(stream)
    w.synchronize(prop)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_events = ev0.elapsed_time(ev1)
    taps = _time_taps(torch, w, prop, physics, n, d, steps, warmup, stream) if with_taps else None
    pc = torch.empty((n, n, n), device=dev)
    hip.hipMemcpy(pc.data_ptr(), w.device_ptr(prop, "pCur"), pc.numel() * 4, 3)
    finite = bool(torch.isfinite(pc).all())
    del pc
    w.free(prop)
    torch.cuda.empty_cache()
    cells = float(n) ** 3
    ms_per_step = wall * 1e3 / steps
    kernel_ms = ms_events / steps
    return {
        "physics": physics, "n": n, "steps": steps, "warmup": warmup, "setup_s": setup_s,
        "ms_per_step": ms_per_step, "kernel_ms": kernel_ms,
        "mcells_s": cells * steps / wall / 1e6,
        "achieved_gbs": cells * BYTES_PER_CELL[physics] / (kernel_ms * 1e-3) / 1e9,
        "finite": finite, "launch": launch, "autotune": tuned, "taps": taps,
    }


# 2D steps: compulsory HBM bytes per cell (Iso: pCur, pOld, v, b in, p+, pSpace out; VTI adds
# mCur, mOld, eps, eta, f in and m+, mSpace out; TTI adds sin/cos theta); Q is the profile and
# comes from its table by sponge distance
BYTES_2D = {"iso": 24, "vti": 52, "tti": 60}


def run_2d(physics, n, steps, warmup):
    """2D forward step on an n x n grid at 12.5 m, dt 1 ms, free surface: HIP-event time per
    step on the handle's stream, Mcells/s and HBM roofline fraction."""
    import torch
    import wavefd_amd.prop2d as w2

    cls = {"iso": w2.Prop2DAcoIsoDenQ_DEO2_FDTD, "vti": w2.Prop2DAcoVTIDenQ_DEO2_FDTD,
           "tti": w2.Prop2DAcoTTIDenQ_DEO2_FDTD}[physics]
    prop = cls(nz=n, nx=n, dz=12.5, dx=12.5, dt=1e-3, nsponge=50, freesurface=True, freqQ=10.0, qMin=0.1,
               qInterior=100.0)
    rng = np.random.default_rng(5)
    z = np.linspace(0.0, 1.0, n, dtype=np.float32)[:, None]
    w2.V(prop)[...] = 1500 + 3000 * z + 50 * rng.standard_normal((n, n)).astype(np.float32)
    w2.B(prop)[...] = 1 / (1.0 + 1.5 * z)
    if physics != "iso":
        w2.Eps(prop)[...] = 0.1
        w2.Eta(prop)[...] = 0.05
        w2.F(prop)[...] = 0.85
    if physics == "tti":
        th = np.deg2rad(30.0) * np.ones((n, n), np.float32)
        w2.SinTheta(prop)[...] = np.sin(th)
        w2.CosTheta(prop)[...] = np.cos(th)
    for g in ((w2.PCur, w2.POld) + ((w2.MCur, w2.MOld) if physics != "iso" else ())):
        g(prop)[...] = rng.standard_normal((n, n)).astype(np.float32) * 1e-3
    w2.sync_to_device(prop)
    w2.coherence(prop, "device")
    stream = torch.cuda.ExternalStream(w2.stream(prop))
    for _ in range(warmup):
        w2.propagateforward(prop)
    w2.synchronize(prop)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        w2.propagateforward(prop)
    ev1.record(stream)
    w2.synchronize(prop)
    ms = ev0.elapsed_time(ev1) / steps
    w2.free(prop)
    cells = float(n) * n
    gbs = cells * BYTES_2D[physics] / (ms * 1e-3) / 1e9
    return {"physics": physics + "2d", "n": n, "steps": steps, "ms_per_step": round(ms, 4),
            "mcells_s": round(cells / (ms * 1e-3) / 1e6, 1), "bytes_per_cell": BYTES_2D[physics],
            "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "kernel": "fused single pass"}


def run_graph(physics="iso", n=96, nt=400, chunk=None):
    """A small window's modelling loop (step, inject, extract per time step): the calls one by
    one against graph segments of `chunk` steps (TimeStepInjectExtractGraph); wall time per step
    with inputs resident on the GPU."""
    import wavefd_amd as w
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[physics]
    f32 = np.float32

    def setup():
        prop = cls(nz=n, ny=n, nx=n, dz=10.0, dy=10.0, dx=10.0, dt=1e-3, nsponge=10, freesurface=True)
        w.V(prop)[...] = 2000
        w.B(prop)[...] = 1
        if physics != "iso":
            w.Eps(prop)[...] = 0.1
            w.Eta(prop)[...] = 0.05
            w.F(prop)[...] = 0.85
        if physics == "tti":
            w.CosTheta(prop)[...] = 1
            w.CosPhi(prop)[...] = 1
        w.sync_to_device(prop)
        w.coherence(prop, "device")
        src = w.hickscoeffs(f32(10), f32(10), f32(10), 0.0, 0.0, 0.0, n, n, n, [100.0], [5.0 * n], [5.0 * n])
        wavelet = w.causal_ricker(15.0, 1e-3 * np.arange(nt))[None, :]
        r = np.linspace(100.0, 10.0 * (n - 10), 32)
        rcv = w.linearcoeffs(f32(10), f32(10), f32(10), 0.0, 0.0, 0.0, n, n, n, np.full(32, 50.0), r, r)
        w.set_sources(prop, w.source_blocking(src, 4), wavelet, fields=3 if physics != "iso" else 1)
        w.set_receivers(prop, w.receiver_blocking(rcv, 4), nt)
        return prop

    prop = setup()
    w.timestep_inject_extract(prop, 1)  # first-call costs outside the timing
    w.synchronize(prop)
    t0 = time.perf_counter()
    for it in range(2, nt + 1):
        w.timestep_inject_extract(prop, it)
    w.synchronize(prop)
    eager = (time.perf_counter() - t0) * 1e3 / (nt - 1)
    d_eager = w.receiver_data(prop, 32, nt)
    w.free(prop)
    prop = setup()
    w.timestep_inject_extract(prop, 1)
    w.synchronize(prop)
    t0 = time.perf_counter()
    it = 2
    chunk = chunk or nt
    while it <= nt:
        k = min(chunk, nt + 1 - it)
        w.timestep_inject_extract_graph(prop, it, k)
        it += k
    graph = (time.perf_counter() - t0) * 1e3 / (nt - 1)
    same = bool(np.array_equal(w.receiver_data(prop, 32, nt), d_eager))
    w.free(prop)
    return {"physics": physics, "n": n, "nt": nt, "chunk": chunk, "eager_ms_per_step": round(eager, 4),
            "graph_ms_per_step": round(graph, 4), "speedup": round(eager / graph, 3), "traces_identical": same}


def run_coherent(physics, n, steps, seed=3):
    """The reference's Julia loop (test/spacetime.jl:486-492) through the default COHERENT host
    API: the model filled through the getters' host arrays, then per step propagateforward
    (host-visible arrays uploaded before and downloaded after), the handle swap, and
    injectdata into the host pcur.  The rate includes every PCIe transfer that keeps host
    arrays coherent."""
    import torch
    import wavefd_amd as w

    dev = torch.device("cuda:0")
    d, dt = 12.5, 1e-3
    cls = {"iso": w.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": w.Prop3DAcoVTIDenQ_DEO2_FDTD,
           "tti": w.Prop3DAcoTTIDenQ_DEO2_FDTD}[physics]
    prop = cls(nz=n, ny=n, nx=n, dz=d, dy=d, dx=d, dt=dt, nsponge=40, freesurface=True, freqQ=10.0,
               qMin=0.1, qInterior=100.0)
    getter = {"v": w.V, "b": w.B, "eps": w.Eps, "eta": w.Eta, "f": w.F}
    if physics == "tti":
        getter.update({"sinTheta": w.SinTheta, "cosTheta": w.CosTheta, "sinPhi": w.SinPhi, "cosPhi": w.CosPhi})
    model = build_model(torch, physics, n, d, dt, seed, dev)
    for name, t in model.items():
        # (nx,ny,nz) C-order == (nz,ny,nx) Julia order
        getter[name](prop)[...] = t.cpu().numpy().transpose(2, 1, 0)
    del model
    torch.cuda.empty_cache()
    pcur, pold = w.PCur(prop), w.POld(prop)
    mid = (n // 2) * d
    blocks = w.source_blocking(w.hickscoeffs(np.float32(d), np.float32(d), np.float32(d), 0.0, 0.0, 0.0,
                                             n, n, n, [20.0], [mid + d / 2], [mid + d / 2]), 8)
    total = steps + 1
    wavelet = w.causal_ricker(15.0, dt * np.arange(total))[None, :].astype(np.float32)
    w.propagateforward(prop)  # first call: host mirrors and staging allocated
    pcur, pold = pold, pcur
    w.injectdata(pcur, blocks, wavelet, 1)
    t0 = time.perf_counter()
    for it in range(2, total + 1):
        w.propagateforward(prop)
        pcur, pold = pold, pcur
        w.injectdata(pcur, blocks, wavelet, it)
    dt_wall = time.perf_counter() - t0
    finite = bool(np.isfinite(pcur).all())
    w.free(prop)
    return {"physics": physics, "n": n, "steps": steps, "mcells_s": float(n) ** 3 * steps / dt_wall / 1e6,
            "ms_per_step": dt_wall * 1e3 / steps, "finite": finite,
            "loop": "coherent host API: propagateforward + swap + host injectdata (test/spacetime.jl:486-492)"}


def run_born(physics, nshots=24, nt=200, seed=4, big=(600, 1000, 1000), fused=True, batch=1,
             snapshot_gb=180.0, checkpoint=None, predictive=None):
    """Born shots/hour: `nshots` shots on seeded ginsu windows (per-axis sizes U{200..600}, from the
    surface down, horizontal origin uniform in the grid) of a big (nz, ny, nx) model at 12.5 m,
    dt 1 ms, nt steps each: background forward saving DP (and DM), Born forward with a seeded
    +-2 % smoothed velocity perturbation, adjoint with the zero-lag imaging condition, window
    image added into the shared image.  Source at 20 m depth in the window centre, a 16 x 16
    receiver grid at 20 m over the window's interior.  fused: the per-step pairs as fused steps;
    False: the separate calls of the reference's API.  batch > 1: up to `batch` shots (and at most
    snapshot_gb of DP snapshots) run together, each time step of all of them one launch;
    batch 1: shots one after another.  checkpoint: recompute DP (DM) from background checkpoints
    every `checkpoint` steps instead of storing all nt (the same image bit for bit); predictive:
    keep them in the predictive store with segments of that many frames (lossy, about 8x fewer
    bytes)."""
    import torch
    from wavefd_amd import born, ginsu

    dev = torch.device("cuda:0")
    d, dt = 12.5, 1e-3
    nz, ny, nx = big
    rng = np.random.default_rng(seed)
    windows = []
    for i in range(nshots):
        r = np.random.default_rng(seed + i)
        sz = [int(r.integers(200, 601)) for _ in range(3)]
        snz, sny, snx = min(sz[0], nz), min(sz[1], ny), min(sz[2], nx)
        oy = int(r.integers(0, ny - sny + 1))
        ox = int(r.integers(0, nx - snx + 1))
        windows.append(ginsu.Window(big, (0, oy, ox), (snz, sny, snx)))
    t0 = time.time()
    hip = _hip()
    # the big model in the bench's construction, moved into buffers of this package's runtime
    # ((nx, ny, nz) C-order == Julia (nz, ny, nx))
    m = build_model(torch, physics, (nx, ny, nz), d, dt, seed + 100, dev)
    bufs = {}
    for name in born.MODEL_FIELDS[physics]:
        t = m[name].contiguous()
        bufs[name] = ginsu.DeviceBuffer(t.numel() * 4)
        torch.cuda.synchronize()
        _copy_d2d(hip, bufs[name].ptr, t)
    dvt = (0.02 * m["v"] * _smooth_noise(torch, (nx, ny, nz), 15, seed + 200, dev)).contiguous()  # +-2 % of v
    torch.cuda.synchronize()
    dvb = ginsu.DeviceBuffer(dvt.numel() * 4)
    _copy_d2d(hip, dvb.ptr, dvt)
    del m, dvt
    torch.cuda.empty_cache()
    image = ginsu.DeviceBuffer(4 * nx * ny * nz)
    image.zero()
    image.sync()
    fields = {k: b.ptr for k, b in bufs.items()}
    setup_s = time.time() - t0
    nf = 1 if physics == "iso" else 2
    maxcells = max(w.ncells for w in windows)
    snap = (ginsu.DeviceBuffer(4 * nt * nf * maxcells) if batch <= 1 and checkpoint is None and predictive is None
            else None)
    wav = np.asarray(born.spacetime.causal_ricker(15.0, dt * np.arange(nt)), np.float32)
    shots = []
    for w in windows:
        oz, oy, ox = w.origin
        snz, sny, snx = w.size
        cy, cx = (oy + sny / 2.0) * d, (ox + snx / 2.0) * d
        ry, rx = np.meshgrid(np.linspace((oy + 30) * d, (oy + sny - 30) * d, 16),
                             np.linspace((ox + 30) * d, (ox + snx - 30) * d, 16), indexing="ij")
        shots.append(born.BornShot((20.0, cy, cx), np.full(ry.size, 20.0), ry.ravel(), rx.ravel(), wav))
    t1 = time.perf_counter()
    if batch > 1:
        born.born_shots(physics, windows, (d, d, d), dt, shots, fields, dvb.ptr, image.ptr, nsponge=30, batch=batch,
                        snapshot_bytes=snapshot_gb * 1e9, keep_born_data=False, fused=fused, checkpoint=checkpoint,
                        predictive=predictive)
    else:
        for w, sh in zip(windows, shots):
            born.born_shot(physics, w, (d, d, d), dt, sh, fields, dvb.ptr, image.ptr, nsponge=30, snapshots=snap,
                           keep_born_data=False, fused=fused, checkpoint=checkpoint, predictive=predictive)
    image.sync()
    wall = time.perf_counter() - t1
    img = image.download(np.zeros((nx, ny, nz), np.float32))
    finite = bool(np.isfinite(img).all()) and float(np.abs(img).max()) > 0
    import hashlib
    image_sha = hashlib.sha256(img.tobytes()).hexdigest()[:16]
    del img
    cells = sum(w.ncells for w in windows)
    if snap is not None:
        snap.free()
    for b in list(bufs.values()) + [dvb, image]:
        b.free()
    return {"physics": physics, "shots": nshots, "nt": nt, "big": list(big), "wall_s": round(wall, 3),
            "image_sha": image_sha,
            "shots_per_hour": round(nshots * 3600.0 / wall, 1),
            "mcell_steps_per_s": round(3 * nt * cells / wall / 1e6, 1), "finite": finite,
            "setup_s": round(setup_s, 2), "window_cells_total": cells,
            "passes": "forward (DP snapshots) + Born forward + adjoint imaging",
            "snapshot_store": ("exact" if checkpoint is None and predictive is None else
                               f"checkpoints every {checkpoint} steps, recomputed" if checkpoint is not None else
                               f"predictive, segments of {predictive} frames (lossy)"),
            "snapshot_gb_largest_window": round(4e-9 * maxcells * (
                born.checkpoint_fields(nt, nf, checkpoint) if checkpoint is not None else
                nt * nf / 8.2 + min(predictive, nt) * nf if predictive is not None else nt * nf), 2),
            "snapshot_gb_largest_window_exact": round(4e-9 * maxcells * nt * nf, 2),
            "shots_at_once": batch if batch > 1 else 1,
            "steps": ("one launch per time step of a batch" if batch > 1 else "fused pairs") if fused
            else "separate calls"}


# the predictive snapshot store's settings: frames per segment, extrapolation order, keyframe bits
PREDICTIVE = (48, 3, 16)


def run_store(physics="vti", n=512, nt=400, seed=5, k=96, levels=(5, 11)):
    """The Born gradient with compressed snapshots judged against exact storage: one shot on an
    n^3 VTI model (the benchmark model's construction) at 12.5 m, dt 1 ms, nt steps, a +-2 % smoothed dv;
    the image with the wave-predicted store (born.py::_WaveStore, segments of k frames, 16-bit
    keyframes, residuals quantised to each of `levels`) against the image with exact snapshots (the
    checkpointed store, which gives the exact store's image bit for bit).  Returns the stored-byte
    ratio against float32, the image's relative L2 error and the wall times, for the first of
    `levels` and, under "variants", for every one."""
    import torch
    from wavefd_amd import born, ginsu

    dev = torch.device("cuda:0")
    d, dt = 12.5, 1e-3
    hip = _hip()
    m = build_model(torch, physics, n, d, dt, seed, dev)
    bufs = {}
    for name in born.MODEL_FIELDS[physics]:
        bufs[name] = ginsu.DeviceBuffer(n ** 3 * 4)
        torch.cuda.synchronize()
        _copy_d2d(hip, bufs[name].ptr, m[name].contiguous())
    dvt = (0.02 * m["v"] * _smooth_noise(torch, (n, n, n), 15, seed + 200, dev)).contiguous()
    torch.cuda.synchronize()
    dvb = ginsu.DeviceBuffer(n ** 3 * 4)
    _copy_d2d(hip, dvb.ptr, dvt)
    del m, dvt
    torch.cuda.empty_cache()
    win = ginsu.Window((n, n, n), (0, 0, 0), (n, n, n))
    ry, rx = np.meshgrid(np.linspace(30 * d, (n - 30) * d, 16), np.linspace(30 * d, (n - 30) * d, 16), indexing="ij")
    wav = np.asarray(born.spacetime.causal_ricker(15.0, dt * np.arange(nt)), np.float32)
    shot = born.BornShot((20.0, n * d / 2, n * d / 2), np.full(ry.size, 20.0), ry.ravel(), rx.ravel(), wav)
    fields = {k_: b.ptr for k_, b in bufs.items()}
    imgs, secs, stored = {}, {}, {}
    for mode in ("exact",) + tuple(levels):
        img = ginsu.DeviceBuffer(n ** 3 * 4)
        img.zero()
        t0 = time.perf_counter()
        r = born.born_shot(physics, win, (d, d, d), dt, shot, fields, dvb.ptr, img.ptr, nsponge=40,
                           keep_born_data=False, checkpoint=20 if mode == "exact" else None,
                           predictive=None if mode == "exact" else (k, "wave", 16, mode))
        img.sync()
        secs[mode] = time.perf_counter() - t0
        if mode != "exact":
            stored[mode] = r["stored_bytes"]
        imgs[mode] = img.download(np.zeros((n, n, n), np.float32)).astype(np.float64)
        img.free()
    for b in list(bufs.values()) + [dvb]:
        b.free()
    nf = 1 if physics == "iso" else 2
    variants = []
    for lev in levels:
        rel = float(np.linalg.norm(imgs[lev] - imgs["exact"]) / np.linalg.norm(imgs["exact"]))
        variants.append({"residual_levels": lev, "stored_ratio_vs_float32": round(4.0 * nf * n ** 3 * nt / stored[lev], 3),
                         "image_rel_l2_vs_exact": rel, "seconds": round(secs[lev], 2)})
    return {"workload": "%s %d^3 at 12.5 m, one Born shot (forward, Born forward, adjoint imaging), nt %d"
                        % (physics.upper(), n, nt),
            "store": "wave-predicted (one propagator step + the source's image as predictor, residuals quantised "
                     "to %d levels, 16-bit keyframes, segments of %d frames)" % (levels[0], k),
            "stored_ratio_vs_float32": variants[0]["stored_ratio_vs_float32"],
            "image_rel_l2_vs_exact": variants[0]["image_rel_l2_vs_exact"],
            "seconds": {"exact_checkpointed": round(secs["exact"], 2), "wave_store": variants[0]["seconds"]},
            "variants": variants,
            "target": ">= 8x fewer snapshot bytes with the Born gradient within rel L2 1e-3 of exact storage"}


def run_codec(n=128, nt=300, iters=800, seed=7, big_n=512):
    """Learned snapshot codec (wavefd_amd/lcodec.py) on the bench's Iso model construction.
    Training: DP snapshots (every 3rd step) of an n^3 shot made by the propagator's fused SaveDP
    steps, decoder trained on the GPU (iters Adam steps of 4 x 32^3 patches).  Evaluation on a
    second model: snapshot errors of the plain and the learned decode, and one Born shot's image
    with exact, plain and learned snapshot storage (the image error judged against exact
    storage).  Rates: encode (with the encoder's choice) and learned decode of a big_n^3 field."""
    import torch
    import wavefd_amd as w
    from wavefd_amd import born, ginsu, lcodec, spacetime as st

    dev = torch.device("cuda:0")
    hip = _hip()
    d, dt = 12.5, 1e-3
    shape = (n, n, n)
    F32 = np.float32

    def model(sd):
        m = build_model(torch, "iso", n, d, dt, sd, dev)
        out = {}
        for k in ("v", "b"):
            t = m[k].contiguous()
            out[k] = ginsu.DeviceBuffer(t.numel() * 4)
            torch.cuda.synchronize()
            _copy_d2d(hip, out[k].ptr, t)
        del m
        return out

    def snapshots(mdl, residual=False):
        """DP every 3rd step; residual: instead the open-loop residual of a third-order temporal
        extrapolation, DP_n - (3 DP_n-1 - 3 DP_n-2 + DP_n-3), what the predictive store codes"""
        p = w.Prop3DAcoIsoDenQ_DEO2_FDTD(nz=n, ny=n, nx=n, dz=d, dy=d, dx=d, dt=dt, nsponge=20, freesurface=True,
                                         freqQ=10.0, qMin=0.1, qInterior=100.0)
        w.coherence(p, "device")
        for k in ("v", "b"):
            ginsu.copy(w.device_ptr(p, k), mdl[k].ptr, 4 * n ** 3)
        ginsu.sync()
        src = st.hickscoeffs(F32(d), F32(d), F32(d), 0.0, 0.0, 0.0, n, n, n, [20.0], [n * d / 2 + 3.0],
                             [n * d / 2 - 4.0], fs_index=1)
        w.set_sources(p, [src], st.causal_ricker(15.0, dt * np.arange(nt)).reshape(1, -1).astype(F32))
        ring = [ginsu.DeviceBuffer(4 * n ** 3) for _ in range(4)]  # DP of the last 4 steps, newest first
        pred = ginsu.DeviceBuffer(4 * n ** 3)
        out = []
        strm = p._lib.GetStream(p.p)
        for it in range(1, nt + 1):
            p._lib.TimeStepSaveDP(p.p, ring[3].ptr)
            p._lib.InjectRegistered(p.p, it)
            ring = [ring[3]] + ring[:3]
            if it % 3 == 0 and (not residual or it >= 4):
                b = ginsu.DeviceBuffer(4 * n ** 3)
                if residual:
                    lcodec.combine(n ** 3, 3.0, ring[1].ptr, pred.ptr, -3.0, ring[2].ptr, 1.0, ring[3].ptr, stream=strm)
                    lcodec.combine(n ** 3, 1.0, ring[0].ptr, b.ptr, -1.0, pred.ptr, stream=strm)
                else:
                    ginsu.copy(b.ptr, ring[0].ptr, 4 * n ** 3, strm)
                out.append(b)
        w.synchronize(p)
        for q in ring + [pred]:
            q.free()
        w.free(p)
        return out

    t0 = time.perf_counter()
    m1 = model(seed)
    train = snapshots(m1)
    t_gen = time.perf_counter() - t0
    codec = lcodec.LearnedCodec(seed=seed)
    t0 = time.perf_counter()
    losses = codec.train([b.ptr for b in train], shape, iters=iters, patch=32, batch=4, lr=1e-3, seed=seed)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    # a second decoder, trained on the residual frames the predictive store codes
    rtrain = snapshots(m1, residual=True)
    rcodec = lcodec.LearnedCodec(seed=seed + 1)
    rlosses = rcodec.train([b.ptr for b in rtrain], shape, iters=iters, patch=32, batch=4, lr=1e-3, seed=seed)
    for b in rtrain:
        b.free()
    for b in train:
        b.free()
    # snapshot errors on the second model
    m2 = model(seed + 1)
    test = snapshots(m2)
    cb = ginsu.DeviceBuffer(4 * lcodec.code_words(shape))
    a, b2, scr = (ginsu.DeviceBuffer(4 * n ** 3) for _ in range(3))
    ep, el = [], []
    for sb in test:
        codec.encode(sb.ptr, shape, cb.ptr, a.ptr, scr.ptr)
        lcodec.dequantize(cb.ptr, a.ptr)
        codec.decode(cb.ptr, b2.ptr, scr.ptr)
        x = sb.download(np.zeros(n ** 3, F32)).astype(np.float64)
        nx_ = np.linalg.norm(x)
        ep.append(np.linalg.norm(a.download(np.zeros(n ** 3, F32)) - x) / nx_)
        el.append(np.linalg.norm(b2.download(np.zeros(n ** 3, F32)) - x) / nx_)
    for q in test + [cb, a, b2, scr]:
        q.free()
    # the Born image with the three snapshot stores
    dvt = (0.02 * build_model(torch, "iso", n, d, dt, seed + 1, dev)["v"] *
           _smooth_noise(torch, shape, 15, seed + 300, dev)).contiguous()
    dvb = ginsu.DeviceBuffer(dvt.numel() * 4)
    torch.cuda.synchronize()
    _copy_d2d(hip, dvb.ptr, dvt)
    del dvt
    win = ginsu.Window(shape, (0, 0, 0), shape)
    ry, rx = np.meshgrid(np.linspace(30 * d, (n - 30) * d, 16), np.linspace(30 * d, (n - 30) * d, 16), indexing="ij")
    shot = born.BornShot((20.0, n * d / 2 + 3.0, n * d / 2 - 4.0), np.full(ry.size, 20.0), ry.ravel(), rx.ravel(),
                         st.causal_ricker(15.0, dt * np.arange(nt)))
    imgs, times, stored = {}, {}, {}
    for mode in ("exact", "plain", "learned", "predictive", "predictive_learned"):
        img = ginsu.DeviceBuffer(4 * n ** 3)
        img.zero()
        t0 = time.perf_counter()
        r = born.born_shot("iso", win, (d, d, d), dt, shot, {k: q.ptr for k, q in m2.items()}, dvb.ptr, img.ptr,
                           nsponge=20, snapshot_codec=codec if mode in ("plain", "learned") else None,
                           learned=mode == "learned", predictive=PREDICTIVE if mode.startswith("predictive") else None,
                           residual_codec=rcodec if mode == "predictive_learned" else None)
        times[mode] = time.perf_counter() - t0
        if r["stored_bytes"]:
            stored[mode] = r["stored_bytes"]
        imgs[mode] = img.download(np.zeros(n ** 3, F32)).astype(np.float64)
        img.free()
    ref_n = np.linalg.norm(imgs["exact"])
    gerr = {k: float(np.linalg.norm(imgs[k] - imgs["exact"]) / ref_n)
            for k in ("plain", "learned", "predictive", "predictive_learned")}
    for q in list(m1.values()) + list(m2.values()) + [dvb]:
        q.free()
    # rates on a big field
    bs = (big_n, big_n, big_n)
    f = _smooth_noise(torch, bs, 4, seed + 500, dev).contiguous()
    fb = ginsu.DeviceBuffer(f.numel() * 4)
    torch.cuda.synchronize()
    _copy_d2d(hip, fb.ptr, f)
    del f
    torch.cuda.empty_cache()
    cb = ginsu.DeviceBuffer(4 * lcodec.code_words(bs))
    o, s1 = ginsu.DeviceBuffer(4 * big_n ** 3), ginsu.DeviceBuffer(4 * big_n ** 3)
    codec.encode(fb.ptr, bs, cb.ptr, o.ptr, s1.ptr)  # warm-up (workspace)
    codec.decode(cb.ptr, o.ptr, s1.ptr)
    ginsu.sync()
    t0 = time.perf_counter()
    codec.encode(fb.ptr, bs, cb.ptr, o.ptr, s1.ptr)
    ginsu.sync()
    t_enc = time.perf_counter() - t0
    t0 = time.perf_counter()
    codec.decode(cb.ptr, o.ptr, s1.ptr)
    ginsu.sync()
    t_dec = time.perf_counter() - t0
    t0 = time.perf_counter()
    lcodec.encode(fb.ptr, bs, cb.ptr)
    lcodec.dequantize(cb.ptr, o.ptr)
    ginsu.sync()
    t_plain = time.perf_counter() - t0
    for q in (fb, cb, o, s1):
        q.free()
    codec.close()
    rcodec.close()
    gb = 4 * big_n ** 3 / 1e9
    return {"ratio_512": round(lcodec.ratio((512, 512, 512)), 3), "ratio_window": round(lcodec.ratio(shape), 3),
            "train": {"snapshots": len(train), "grid": n, "iters": iters, "patch": 32, "batch": 4, "seconds":
                      round(t_train, 2), "loss_first": float(losses[:20].mean()),
                      "loss_last": float(losses[-50:].mean()), "generate_seconds": round(t_gen, 2)},
            "snapshot_rel_err": {"plain": round(float(np.mean(ep)), 4), "learned": round(float(np.mean(el)), 4),
                                 "snapshots": len(ep)},
            "born_image_rel_err_vs_exact": {k: round(v, 6) for k, v in gerr.items()},
            "predictive": {"segment": PREDICTIVE[0], "order": PREDICTIVE[1], "keyframe_bits": PREDICTIVE[2],
                           "stored_ratio_vs_exact": round(4 * n ** 3 * nt / stored["predictive"], 3),
                           "residuals": "11-level block quantiser; predictive_learned: decoded by a second "
                                        "network trained on open-loop residual frames",
                           "residual_train_loss_first_last": [float(rlosses[:20].mean()), float(rlosses[-50:].mean())]},
            "born_seconds": {k: round(v, 2) for k, v in times.items()},
            "rates_GBps_512": {"encode_learned": round(gb / t_enc, 1), "decode_learned": round(gb / t_dec, 1),
                               "encode_plus_dequantize_plain": round(gb / t_plain, 1)},
            "target": ">= 8x fewer snapshot bytes with the Born gradient within rel L2 1e-3 of exact storage"}


def _host_info():
    """nproc, the CPUs this process may run on, and lscpu's model/socket/core lines"""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
                               "NUMA node(s)"):
                info["lscpu " + key.strip()] = val.strip()
    except (OSError, subprocess.SubprocessError):
        info["lscpu"] = "unavailable"
    return info


def _cpu_child(physics, n, target_s):
    """(child process) the oracle's two-pass step on the benchmark's n^3 model, FTZ/DAZ on"""
    host = _host_info()  # before OpenMP binds this thread to one place
    from oracle import oracle as orc
    from tests.helpers import config_angles, config_model, config_q, smooth_noise_big

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    cls = {"iso": orc.OracleIso, "vti": orc.OracleVTI, "tti": orc.OracleTTI}[physics]
    d, dt = 12.5, 1e-3
    nsponge = 50 if physics == "tti" else 40
    t0 = time.perf_counter()
    prop = cls(n, n, n, d, d, d, dt=dt, fs=True, nsponge=nsponge, freqQ=10.0, qMin=0.1, qInterior=100.0,
               nthread=threads)
    m = config_model(physics, (n, n, n), d, dt, 3)
    for name, arr in m.items():
        getattr(prop, name)[...] = arr
    del m
    if physics == "tti":
        # tilt in [-45, 45] deg and azimuth in [0, 360) deg, smoothed by 25 cells (azimuth as a unit
        # vector); Q from 30 to 200 in the interior, the sponge keeping SetupDtOmegaInvQ's profile
        for name, arr in config_angles((n, n, n), 303).items():
            getattr(prop, name)[...] = arr
        config_q(prop.q, nsponge, dt, 306)
    p0 = smooth_noise_big((n, n, n), 4, 307)
    for name in ("pCur", "pOld") + (("mCur", "mOld") if physics != "iso" else ()):
        getattr(prop, name)[...] = p0
    del p0
    setup_s = time.perf_counter() - t0
    orc.lib().wo_set_ftz(threads, 1)
    prop.time_step()  # warm-up (first touch of the temporaries)
    t0 = time.perf_counter()
    k = 0
    while True:
        prop.time_step()
        k += 1
        if time.perf_counter() - t0 >= target_s:
            break
    el = time.perf_counter() - t0
    return {"value": n ** 3 * k / el / 1e6, "unit": "Mcells/s", "cores": threads, "kind": "port",
            "sample": "oracle two-pass %s step (the reference's structure, nbx = nby = 8, nbz = nz) on the "
                      "benchmark's %d^3 model, %d steps in %.1f s after one warm-up step" % (physics.upper(), n, k, el),
            "settings": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
                         "OMP_PLACES": os.environ.get("OMP_PLACES"), "ftz_daz": True,
                         "model_setup_s": round(setup_s, 1)},
            "host": host}


def cpu_baseline(physics, n, target_s=12.0):
    """Time the oracle's two-pass step (the reference's structure) on the host cores for about
    target_s seconds on the benchmark's n^3 model, in a child process started with OMP_PROC_BIND=close
    and OMP_PLACES=cores (benchmark/README.md:8) and flush-to-zero on; returns its JSON record."""
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", "--physics", physics, "--cpu-n", str(n),
           "--cpu-seconds", str(target_s)]
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        raise RuntimeError("CPU baseline failed:\n" + out.stderr[-2000:])
    return json.loads(out.stdout.strip().splitlines()[-1])


# the sources the step kernels are built from: a traffic record applies to the build it measured
KERNEL_SOURCES = ("wfd_fused.hpp", "wfd_wide.hpp", "wfd_kernels.hpp", "wfd_runtime.hpp")


def kernel_source_hash():
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "wavefd_amd", "csrc", name), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(physics, n):
    """Bytes per launch of the fused step kernel between the L2 and the memory fabric (so
    Infinity-Cache hits included: an upper bound on HBM bytes) from the committed rocprofv3 --pmc
    runs (profiles/k_wide_step_traffic.json: FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each
    scaled by what the counter reads for a plain copy kernel of known size), or None when no
    measurement of this workload on this build of the kernels is committed."""
    path = os.path.join(ROOT, "profiles", "k_wide_step_traffic.json")
    try:
        with open(path) as fh:
            doc = json.load(fh)
    except (OSError, ValueError):
        return None
    rec = doc.get("%s_%d" % (physics, n))
    if not rec or doc.get("kernel_source_hash") != kernel_source_hash():
        return None
    return rec["fetch_kib"] * rec["fetch_scale"] * 1024 + rec["write_kib"] * rec["write_scale"] * 1024


def copy_probe(nbytes=1 << 30, reps=3):
    """A plain copy kernel of known size: reads and writes nbytes per launch."""
    import torch
    src = torch.rand(nbytes // 4, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--physics", default="tti", choices=["iso", "vti", "tti"])
    ap.add_argument("--n", type=int, default=800)
    ap.add_argument("--cpu-n", type=int, default=800)
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the Iso 256^3 and VTI 512^3 steps")
    ap.add_argument("--no-born", action="store_true",
                    help="skip the Born shots/hour (fused and separate calls, Iso and VTI, 24 shots each) and "
                         "the compressed-snapshot gradient measurements of the default line")
    ap.add_argument("--born", choices=["iso", "vti"], default=None,
                    help="only measure Born shots/hour on ginsu windows of a 1000x1000x600 model")
    ap.add_argument("--born-shots", type=int, default=24)
    ap.add_argument("--born-nt", type=int, default=200)
    ap.add_argument("--born-plain", action="store_true", help="Born with the separate per-step calls")
    ap.add_argument("--born-batch", type=int, default=1, help="Born shots run together (one launch per step)")
    ap.add_argument("--born-predictive", type=int, default=None,
                    help="Born: keep DP in the predictive store with segments of K frames (lossy, ~8x fewer bytes)")
    ap.add_argument("--born-checkpoint", type=int, default=None,
                    help="Born: recompute DP from background checkpoints every K steps instead of storing them")
    ap.add_argument("--codec", action="store_true", help="only train and measure the learned snapshot codec")
    ap.add_argument("--coherent", action="store_true",
                    help="also time the reference's host loop through the coherent API (PCIe-inclusive)")
    ap.add_argument("--graph", action="store_true",
                    help="only time a small window's modelling loop, calls one by one against graph segments")
    ap.add_argument("--prop2d", action="store_true",
                    help="only time the 2D Iso, VTI and TTI steps (8192^2 cells)")
    ap.add_argument("--autotune", action="store_true",
                    help="choose the fused step's launch configuration with the autotuner before timing "
                         "(measured: it picks the automatic rule's configuration; off by default so that a "
                         "kernel-trace of the default command holds only the timed configuration)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy kernel 3 times (the known-size "
                         "kernel that calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.cpu_child:
        print(json.dumps(_cpu_child(args.physics, args.cpu_n, args.cpu_seconds)), flush=True)
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.codec:
        print(json.dumps({"codec": run_codec()}), flush=True)
        return
    if args.graph:
        print(json.dumps({"graph": [run_graph(ph, n) for ph in ("iso", "tti") for n in (64, 128)]}), flush=True)
        return
    if args.prop2d:
        print(json.dumps({"prop2d": [run_2d(ph, 8192, 50, 5) for ph in ("iso", "vti", "tti")]}), flush=True)
        return
    if args.born:
        print(json.dumps({"born": run_born(args.born, args.born_shots, args.born_nt, fused=not args.born_plain,
                                           batch=args.born_batch, checkpoint=args.born_checkpoint,
                                           predictive=args.born_predictive)}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the propagator state of one shot lives on one MI355X and this "
              "build has no multi-GPU path; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    def progress(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    res = run_gpu(args.physics, args.n, args.steps, args.warmup, tune=args.autotune, with_taps=True)
    progress("%s %d^3 step: %.3f ms" % (args.physics, args.n, res["kernel_ms"]))
    extra = []
    if not args.no_extra:
        for ph, n in (("iso", 256), ("vti", 512)):
            r = run_gpu(ph, n, max(10, args.steps), args.warmup)
            progress("%s %d^3 step: %.4f ms" % (ph, n, r["kernel_ms"]))
            extra.append(dict({k: r[k] for k in ("physics", "n", "mcells_s", "kernel_ms", "achieved_gbs")},
                              frac=round(r["achieved_gbs"] / HBM_PEAK_GBS, 4), bytes_per_cell=BYTES_PER_CELL[ph],
                              launch=r["launch"]))
    born_line = None
    if not args.no_born:
        born_line = {}
        for ph, nt in (("iso", 200), ("vti", 100)):
            fused = run_born(ph, 24, nt, fused=True)
            progress("born %s fused: %.1f shots/h" % (ph, fused["shots_per_hour"]))
            plain = run_born(ph, 24, nt, fused=False)
            progress("born %s separate calls: %.1f shots/h" % (ph, plain["shots_per_hour"]))
            born_line[ph] = {"shots": 24, "nt": nt, "shots_per_hour_fused": fused["shots_per_hour"],
                             "shots_per_hour_separate_calls": plain["shots_per_hour"],
                             "ratio": round(fused["shots_per_hour"] / plain["shots_per_hour"], 3),
                             "wall_s_fused": fused["wall_s"], "wall_s_separate_calls": plain["wall_s"],
                             "image_rel_l2_fused_vs_separate": 0.0 if fused["image_sha"] == plain["image_sha"]
                             else None}
        born_line["workload"] = ("24 shots on seeded ginsu windows (per-axis sizes U{200..600}) of a 1000 x 1000 x "
                                 "600 model at 12.5 m: forward (DP snapshots, exact) + Born forward + adjoint "
                                 "imaging into the shared image; fused: the per-step pairs as fused steps (Iso: "
                                 "background and Born field in one pass); separate calls: the reference's API "
                                 "calls one by one")
        store = run_store()
        progress("snapshot store: %.2fx, image rel L2 %.2e" % (store["stored_ratio_vs_float32"],
                                                               store["image_rel_l2_vs_exact"]))
        born_line["snapshot_compression"] = store
    if args.coherent:
        for ph, n in (("iso", 256), (args.physics, min(args.n, 512))):
            extra.append({"coherent": run_coherent(ph, n, 5)})
    cpu = None if args.no_cpu else cpu_baseline(args.physics, args.cpu_n, args.cpu_seconds)
    frac = res["achieved_gbs"] / HBM_PEAK_GBS
    line = {
        "metric": "Mcells/s (3D Iso/TTI DenQ fwd), % HBM roofline; Born shots/h at fixed grad err",
        "value": round(res["mcells_s"], 2),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(res["mcells_s"] / REFERENCE_44T_MCELLS[args.physics], 3),
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3D Aco%sDenQ_DEO2_FDTD forward step, %d^3 at 12.5 m, dt 1 ms, free surface, "
                        "fused single-pass kernel" % (args.physics.upper() if args.physics != "iso" else "Iso", args.n),
            "cells": args.n ** 3,
            "vs_baseline_basis": "reference's published 44-thread CPU Mcells/s for this step "
                                 "(docs/src/index.md:42-44)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(res["achieved_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(frac, 4),
            "traffic": measured_traffic(args.physics, args.n),
            "bytes_per_cell": BYTES_PER_CELL[args.physics],
            "kernel_ms": round(res["kernel_ms"], 4),
            "launch": res["launch"],
        },
        "cpu_baseline": cpu,
    }
    if res["taps"]:
        # the step with registered source injection and receiver extraction, timed separately
        t = res["taps"]
        line["step_inject_extract"] = dict(t, ms_per_step=round(t["ms_per_step"], 4),
                                           gpu_ms_per_step=round(t["gpu_ms_per_step"], 4),
                                           overhead_vs_step=round(t["gpu_ms_per_step"] / res["kernel_ms"] - 1, 4))
    if res["autotune"]:
        line["autotune"] = res["autotune"]
    if extra:
        line["extra"] = extra
    if born_line:
        line["born"] = born_line
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
