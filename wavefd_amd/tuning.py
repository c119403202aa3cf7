"""Launch-configuration autotuner of the fused 3D step -- the GPU analogue of the reference's
cache-block autotuner `tilesize` (src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:159-185, and its VTI/TTI
twins), which times `propagateforward!` over a grid of (nby, nbx) CPU blocks.

On the GPU the blocking that matters is the fused kernel's: its tile shape (64 z-lanes x 16
y-rows or 32 x 32 threads per workgroup), whether the workgroups of an XCD run in rounds kept
in step (shared halo lines in L2), and the length of the x-chunks a workgroup marches over.
None of them changes the results (every configuration evaluates the same expressions in the
same order).  `autotune` times the candidates on the caller's own handle -- its model, its
size -- with HIP events on the handle's stream, restores the wavefields it stepped bit for bit,
and keeps the fastest configuration on the handle.

`tilesize` keeps the reference's signature and return value; the GPU kernels do not use the
CPU block sizes it sweeps, so its table is flat and `autotune` is the tuner that selects
something.
"""
import ctypes

import numpy as np

from . import prop3d as _p3

_WAVE = ("pCur", "pOld", "mCur", "mOld", "pSpace", "mSpace")


class _Hip:
    """The few HIP runtime calls the tuner needs (events on a given stream, device copies)."""

    def __init__(self):
        lib = None
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if lib is None:
            raise RuntimeError("autotune: the HIP runtime (libamdhip64.so) is not loadable")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        for fn, args in (("hipEventCreate", [ctypes.POINTER(vp)]), ("hipEventDestroy", [vp]),
                         ("hipEventRecord", [vp, vp]), ("hipEventSynchronize", [vp]),
                         ("hipEventElapsedTime", [ctypes.POINTER(ctypes.c_float), vp, vp]),
                         ("hipMalloc", [ctypes.POINTER(vp), sz]), ("hipFree", [vp]),
                         ("hipMemcpy", [vp, vp, sz, ctypes.c_int]), ("hipDeviceSynchronize", [])):
            f = getattr(lib, fn)
            f.argtypes = args
            f.restype = ctypes.c_int
        self.lib = lib

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError("autotune: %s failed (%d)" % (what, rc))

    def event(self):
        e = ctypes.c_void_p()
        self.check(self.lib.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e

    def malloc(self, nbytes):
        d = ctypes.c_void_p()
        self.check(self.lib.hipMalloc(ctypes.byref(d), nbytes), "hipMalloc")
        return d

    def copy_d2d(self, dst, src, nbytes):
        self.check(self.lib.hipMemcpy(dst, src, nbytes, 3), "hipMemcpy")


def fused_config(prop):
    """What a plain TimeStep launches with: {"tile", "lockstep", "chunk", "workgroups"}."""
    out = (ctypes.c_long * 4)()
    if prop._lib.GetFusedConfig(prop.p, out) != 0:
        prop._check()
    tile = {1: "64x16", 2: "32x32", 3: "wide", 4: "wide2"}[out[0]]
    return {"tile": tile, "lockstep": bool(out[1]), "chunk": int(out[2]), "workgroups": int(out[3])}


def set_fused_config(prop, tile="auto", lockstep=None, chunk=0):
    """Set the fused step's launch configuration: tile "auto" | "64x16" | "32x32" | "wide" | "wide2", lockstep
    None (automatic) | False | True, chunk 0 (automatic) or x-planes per chunk."""
    ls = -1 if lockstep is None else (1 if lockstep else 0)
    if prop._lib.SetFusedConfig(prop.p, _p3.FUSED_TILES[tile], ls, int(chunk)) != 0:
        prop._check()


def _candidates(prop):
    """Distinct effective configurations: every tile, rounds in step on and off, and the
    x-chunk lengths that split nx into 1, 2, 3, 4, 5, 6, 8, 10 or 12 chunks besides the automatic
    one (rounds in step take effect only where a split gives 20 rounds of workgroups, 12 on the wide
    tile)."""
    nx = prop.shape[2]
    seen, out = set(), []
    for tile in ("64x16", "32x32", "wide", "wide2"):
        for lockstep in (False, True):
            for nxc in (0, 1, 2, 3, 4, 5, 6, 8, 10, 12):
                chunk = 0 if nxc == 0 else -(-nx // nxc)
                set_fused_config(prop, tile, lockstep, chunk)
                eff = fused_config(prop)
                key = (eff["tile"], eff["lockstep"], eff["chunk"])
                if key not in seen:
                    seen.add(key)
                    out.append(key)
    return out


def autotune(prop, nsteps=6, warmup=2, verbose=False):
    """Time every candidate launch configuration of the fused step on `prop` and keep the
    fastest on it.  Returns (best, table): best = {"tile", "lockstep", "chunk", "ms"} and table
    a list of (tile, lockstep, chunk, ms per step).  The wavefields (pCur, pOld, mCur, mOld,
    pSpace, mSpace) are restored bit for bit afterwards, so a tuned handle continues where it
    was; the host-coherence mode is restored too.  Each candidate runs an even number of steps
    (warmup is raised by one if needed), so the pCur/pOld (mCur/mOld) buffers end where they
    began and host views the caller holds keep their names."""
    if not isinstance(prop, _p3._Prop3D):
        raise TypeError("autotune tunes the 3D fused step (Prop3DAco*DenQ_DEO2_FDTD)")
    if (warmup + nsteps) % 2:
        warmup += 1
    hip = _Hip()
    mode = _p3.coherence(prop)
    _p3.sync_to_device(prop)
    _p3.coherence(prop, "device")
    names = [n for n in _WAVE if prop.KIND != "iso" or n[0] == "p"]
    nbytes = prop.n * 4
    saved = {}
    ev0, ev1 = hip.event(), hip.event()
    table = []
    try:
        _p3.synchronize(prop)
        for n in names:
            saved[n] = hip.malloc(nbytes)
            hip.copy_d2d(saved[n], _p3.device_ptr(prop, n), nbytes)
        stream = ctypes.c_void_p(prop._lib.GetStream(prop.p))
        for tile, lockstep, chunk in _candidates(prop):
            set_fused_config(prop, tile, lockstep, chunk)
            for _ in range(warmup):
                _p3.propagateforward(prop)
            hip.check(hip.lib.hipEventRecord(ev0, stream), "hipEventRecord")
            for _ in range(nsteps):
                _p3.propagateforward(prop)
            hip.check(hip.lib.hipEventRecord(ev1, stream), "hipEventRecord")
            hip.check(hip.lib.hipEventSynchronize(ev1), "hipEventSynchronize")
            ms = ctypes.c_float()
            hip.check(hip.lib.hipEventElapsedTime(ctypes.byref(ms), ev0, ev1), "hipEventElapsedTime")
            table.append((tile, lockstep, chunk, ms.value / nsteps))
            if verbose:
                print("autotune: tile %s lockstep %d chunk %d: %.4f ms" % (tile, lockstep, chunk, ms.value / nsteps))
        tile, lockstep, chunk, ms = min(table, key=lambda r: r[3])
        set_fused_config(prop, tile, lockstep, chunk)
        best = {"tile": tile, "lockstep": lockstep, "chunk": chunk, "ms": ms}
    finally:
        _p3.synchronize(prop)
        # the swaps of the timed steps permute which buffer is pCur; the device addresses are
        # read again so each saved field lands under its own name
        for n, d in saved.items():
            hip.copy_d2d(_p3.device_ptr(prop, n), d, nbytes)
            hip.check(hip.lib.hipFree(d), "hipFree")
        hip.lib.hipEventDestroy(ev0)
        hip.lib.hipEventDestroy(ev1)
        if mode == "coherent":
            _p3.sync_to_host(prop)
            _p3.coherence(prop, "coherent")
    return best, table


def tilesize(cls, *, nz, ny, nx, rngy=range(4, 65, 2), rngx=range(4, 65, 2), N=10, nthreads=None):
    """The reference's `tilesize` (src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:159-185): for each (by, bx)
    build a propagator with nby = by, nbx = bx, v = 1500, b = 1 (and eps = eta = 0, f = 0.85,
    the identity rotation for TTI), random wavefields, time N propagateforward calls and
    return (rngx, rngy, speeds) with speeds[iby, ibx] in Mcells/s.  The GPU kernels do not use
    these CPU block sizes; `autotune` selects the GPU's blocking."""
    import time
    rngy, rngx = list(rngy), list(rngx)
    speeds = np.zeros((len(rngy), len(rngx)))
    rng = np.random.default_rng(0)
    for iby, by in enumerate(rngy):
        for ibx, bx in enumerate(rngx):
            prop = cls(nz=nz, ny=ny, nx=nx, nbz=nz, nby=by, nbx=bx, nthreads=nthreads)
            _p3.V(prop)[...] = 1500
            _p3.B(prop)[...] = 1
            if prop.KIND != "iso":
                _p3.F(prop)[...] = 0.85
            if prop.KIND == "tti":
                _p3.CosTheta(prop)[...] = 1
                _p3.CosPhi(prop)[...] = 1
            for g in (_p3.PCur, _p3.POld):
                g(prop)[...] = rng.random(g(prop).shape, dtype=np.float32)
            t = 0.0
            for _ in range(N):
                t0 = time.perf_counter()
                _p3.propagateforward(prop)
                t += time.perf_counter() - t0
            speeds[iby, ibx] = N * prop.n / 1e6 / t
            _p3.free(prop)
    return rngx, rngy, speeds


__all__ = ["autotune", "tilesize", "fused_config", "set_fused_config"]
