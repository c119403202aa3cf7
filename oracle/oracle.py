"""CPU oracle for the WaveFD 3D acoustic path -- TEST INFRASTRUCTURE ONLY.

Tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` compare the
HIP product path against it.  The product package
(``wavefd_amd``) never imports it.

It wraps ``oracle/lib/libwavefd_oracle.so`` (a C restatement of the reference kernels,
see ``wavefd_oracle.c``) and composes the Born / imaging-condition methods of the
reference classes from those primitives with float32 numpy arithmetic, in the reference's
expression order.  Arrays are C-ordered ``(nx, ny, nz)`` float32, i.e. the reference's
``k = kx*ny*nz + ky*nz + kz`` layout (Julia's ``(nz, ny, nx)`` column-major view).

Function-level pins are listed in the header of ``wavefd_oracle.c`` and exercised by
``tests/test_oracle_*.py``.  ``wavefield_separation`` uses numpy's FFT (pocketfft) in
float64; the reference uses FFTW in float32, whose bits are not available here, so that
function is pinned only to the reference's sign/order heuristics
(test/imgcondition.jl:122-129) and to a closed form for a single plane wave.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libwavefd_oracle.so")
_REF_SPACETIME = os.path.join(_HERE, "_ref", "libspacetime_ref.so")

_lib = None

F32 = np.float32
C8 = (F32(1225.0 / 1024.0), F32(-245.0 / 3072.0), F32(49.0 / 5120.0), F32(-5.0 / 7168.0))


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def build_ref():
    """Compile the reference's own src/spacetime.cc into oracle/_ref (needs /root/reference)."""
    subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def ref_spacetime():
    """ctypes handle on the reference's compiled spacetime.cc, or None when absent."""
    if not os.path.exists(_REF_SPACETIME):
        return None
    return ctypes.CDLL(_REF_SPACETIME)


_fp = ctypes.POINTER(ctypes.c_float)
_L = ctypes.c_long
_F = ctypes.c_float


def _declare(l):
    l.wo_setup_q.argtypes = [_L, _L, _L, _L, _L, _F, _F, _F, _F, _fp]
    l.wo_setup_q.restype = ctypes.c_int
    d3 = [_L, _L, _L, _L, _F, _F, _F, _fp, _fp, _fp, _fp, _fp, _fp, _L, _L, _L, _L]
    l.wo_dplus.argtypes = d3
    l.wo_dminus.argtypes = d3
    l.wo_iso_sandwich.argtypes = [_L, _L, _L, _L, _F, _F, _F, _fp, _fp, _fp, _fp, _fp, _L, _L, _L, _L]
    l.wo_iso_update.argtypes = [_L, _L, _L, _L, _F, _F, _F, _F] + [_fp] * 9 + [_L, _L, _L, _L]
    l.wo_vti_sandwich.argtypes = [_L, _L, _L, _L, _F, _F, _F] + [_fp] * 12 + [_L, _L, _L, _L]
    l.wo_vti_update.argtypes = [_L, _L, _L, _L, _F, _F, _F, _F] + [_fp] * 15 + [_L, _L, _L, _L]
    l.wo_tti_sandwich.argtypes = [_L, _L, _L, _L, _F, _F, _F] + [_fp] * 16 + [_L, _L, _L, _L]
    rot = [_L, _L, _L, _L, _F, _F, _F] + [_fp] * 10 + [_L, _L, _L, _L]
    l.wo_tti_dplus_rot.argtypes = rot
    l.wo_tti_dminus_rot.argtypes = rot
    l.wo_iso_step.argtypes = [_L, _L, _L, _L, _F, _F, _F, _F] + [_fp] * 7 + [_L, _L, _L, _L]
    l.wo_vti_step.argtypes = [_L, _L, _L, _L, _F, _F, _F, _F] + [_fp] * 13 + [_L, _L, _L, _L]
    l.wo_tti_step.argtypes = [_L, _L, _L, _L, _F, _F, _F, _F] + [_fp] * 17 + [_L, _L, _L, _L]
    l.wo_inject.argtypes = [_fp, _fp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p, _L]
    l.wo_extract.argtypes = [_fp, _fp, ctypes.c_void_p, _L, ctypes.c_size_t, ctypes.c_size_t]
    l.wo_interp_forward.argtypes = [_fp, _fp, _fp, _L, _L, _L, _L]
    l.wo_interp_adjoint.argtypes = [_fp, _fp, _fp, _L, _L, _L, _L]
    l.wo_max_threads.restype = ctypes.c_int
    l.wo_set_ftz.argtypes = [_L, ctypes.c_int]


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_fp)


def inv(d):
    """1/d evaluated in double and stored to float (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:68-70)."""
    return F32(1.0 / float(F32(d)))


def nthreads_default():
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1


# --------------------------------------------------------------------------------------
# primitives
# --------------------------------------------------------------------------------------

def setup_q(fs, nx, ny, nz, nsponge, dt, freqQ, qMin, qInterior, out=None):
    """src/absorb.h:85-151.  Returns the array, or raises where the reference exits."""
    q = np.zeros((nx, ny, nz), F32) if out is None else out
    rc = lib().wo_setup_q(int(fs), nx, ny, nz, nsponge, F32(dt), F32(freqQ), F32(qMin), F32(qInterior), _p(q))
    if rc != 0:
        raise ValueError("freqQ [%f] is too small" % freqQ)
    return q


class Grid:
    def __init__(self, nx, ny, nz, dx, dy, dz, dt, fs, nbx=8, nby=8, nbz=None, nthread=None):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy, self.dz, self.dt = F32(dx), F32(dy), F32(dz), F32(dt)
        self.fs = 1 if fs else 0
        self.idx, self.idy, self.idz = inv(dx), inv(dy), inv(dz)
        self.nbx, self.nby = nbx, nby
        self.nbz = self.nz if nbz is None else nbz
        self.nthread = nthread or nthreads_default()

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

    def head(self):
        return (self.fs, self.nx, self.ny, self.nz, self.idx, self.idy, self.idz)

    def tail(self):
        return (self.nbx, self.nby, self.nbz, self.nthread)

    def zeros(self, n=1):
        if n == 1:
            return np.zeros(self.shape, F32)
        return [np.zeros(self.shape, F32) for _ in range(n)]


def dplus(g, inX, inY, inZ):
    """src/propagatorStaticFunctions.h:317-528"""
    o = g.zeros(3)
    lib().wo_dplus(*g.head(), _p(inX), _p(inY), _p(inZ), _p(o[0]), _p(o[1]), _p(o[2]), *g.tail())
    return o


def dminus(g, inX, inY, inZ):
    """src/propagatorStaticFunctions.h:534-738"""
    o = g.zeros(3)
    lib().wo_dminus(*g.head(), _p(inX), _p(inY), _p(inZ), _p(o[0]), _p(o[1]), _p(o[2]), *g.tail())
    return o


def tti_dplus_rot(g, g1, g2, g3, sT, cT, sP, cP):
    """src/prop3DAcoTTIDenQ_DEO2_FDTD.h:2910-3165"""
    o = g.zeros(3)
    lib().wo_tti_dplus_rot(*g.head(), _p(g1), _p(g2), _p(g3), _p(sT), _p(cT), _p(sP), _p(cP),
                           _p(o[0]), _p(o[1]), _p(o[2]), *g.tail())
    return o


def tti_dminus_rot(g, g1, g2, g3, sT, cT, sP, cP):
    """src/prop3DAcoTTIDenQ_DEO2_FDTD.h:3171-3657"""
    o = g.zeros(3)
    lib().wo_tti_dminus_rot(*g.head(), _p(g1), _p(g2), _p(g3), _p(sT), _p(cT), _p(sP), _p(cP),
                            _p(o[0]), _p(o[1]), _p(o[2]), *g.tail())
    return o


def iso_sandwich(g, p, b):
    o = g.zeros(3)
    lib().wo_iso_sandwich(*g.head(), _p(p), _p(b), _p(o[0]), _p(o[1]), _p(o[2]), *g.tail())
    return o


def vti_sandwich(g, p, m, eps, eta, f, b):
    o = g.zeros(6)
    lib().wo_vti_sandwich(*g.head(), _p(p), _p(m), _p(eps), _p(eta), _p(f), _p(b), *[_p(a) for a in o], *g.tail())
    return o


def tti_sandwich(g, p, m, sT, cT, sP, cP, eps, eta, f, b):
    o = g.zeros(6)
    lib().wo_tti_sandwich(*g.head(), _p(p), _p(m), _p(sT), _p(cT), _p(sP), _p(cP), _p(eps), _p(eta),
                          _p(f), _p(b), *[_p(a) for a in o], *g.tail())
    return o


# --------------------------------------------------------------------------------------
# propagator state mirrors (reference classes)
# --------------------------------------------------------------------------------------

class _OracleProp:
    """Common state: the reference's arrays, zero-initialised
    (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:73-88), and pointer swap after each step."""

    MODEL = ()
    FIELDS = ()

    def __init__(self, nx, ny, nz, dx, dy, dz, dt, fs=True, nsponge=60, freqQ=5.0, qMin=0.1,
                 qInterior=100.0, nthread=None, setup_q_now=True):
        self.g = Grid(nx, ny, nz, dx, dy, dz, dt, fs, nthread=nthread)
        self.nsponge = nsponge
        for name in self.MODEL + self.FIELDS + ("q",):
            setattr(self, name, self.g.zeros())
        if setup_q_now:
            self.setup_q(freqQ, qMin, qInterior)

    def setup_q(self, freqQ, qMin, qInterior):
        g = self.g
        setup_q(g.fs, g.nx, g.ny, g.nz, self.nsponge, g.dt, freqQ, qMin, qInterior, out=self.q)

    def scale_spatial_derivatives(self):
        """e.g. src/prop3DAcoIsoDenQ_DEO2_FDTD.h:269-291 (full grid)"""
        with np.errstate(all="ignore"):
            s = self.v * self.v / self.b
            for name in self.SPACE:
                getattr(self, name)[...] *= s

    def _dt2v2OverB(self):
        dt = self.g.dt
        return dt * dt * self.v * self.v / self.b


class OracleIso(_OracleProp):
    """Prop3DAcoIsoDenQ_DEO2_FDTD (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:16-1286)"""

    MODEL = ("v", "b")
    FIELDS = ("pCur", "pOld", "pSpace")
    SPACE = ("pSpace",)

    def time_step(self):
        """src/prop3DAcoIsoDenQ_DEO2_FDTD.h:247-261"""
        g = self.g
        tmp = np.empty(3 * self.v.size, F32)
        lib().wo_iso_step(*g.head(), g.dt, _p(self.v), _p(self.b), _p(self.q), _p(self.pCur),
                          _p(self.pSpace), _p(self.pOld), _p(tmp), *g.tail())
        self.pCur, self.pOld = self.pOld, self.pCur

    # Born injection: src/prop3DAcoIsoDenQ_DEO2_FDTD.h:305-498
    def forward_born_injection_V(self, dV, DP):
        with np.errstate(all="ignore"):
            V, B = self.v, self.b
            dt2v2OverB = self._dt2v2OverB()
            factorV = 2 * B * dV / (V * V * V)
            self.pCur += dt2v2OverB * DP * factorV

    def _born_vb(self, dV, dB, P, DP, use_v):
        g = self.g
        t1 = dplus(g, P, P, P)
        t2 = [dB * t for t in t1]
        t1 = dminus(g, *t2)
        with np.errstate(all="ignore"):
            V, B = self.v, self.b
            dt2v2OverB = self._dt2v2OverB()
            factorB = -dB / (V * V)
            if use_v:
                factorV = 2 * B * dV / (V * V * V)
                self.pCur += dt2v2OverB * (DP * (factorV + factorB) + t1[0] + t1[1] + t1[2])
            else:
                self.pCur += dt2v2OverB * (DP * factorB + t1[0] + t1[1] + t1[2])

    def forward_born_injection_VB(self, dV, dB, P, DP):
        self._born_vb(dV, dB, P, DP, True)

    def forward_born_injection_B(self, dB, P, DP):
        self._born_vb(None, dB, P, DP, False)

    # Born imaging: src/prop3DAcoIsoDenQ_DEO2_FDTD.h:518-656
    def adjoint_born_accumulation_V(self, gV, DP):
        with np.errstate(all="ignore"):
            factorV = 2 * self.b / (self.v * self.v * self.v)
            gV += factorV * DP * self.pOld

    def _acc_vb(self, gV, gB, P, DP):
        g = self.g
        t1 = dplus(g, P, P, P)
        t2 = dplus(g, self.pOld, self.pOld, self.pOld)
        with np.errstate(all="ignore"):
            V, B = self.v, self.b
            if gV is not None:
                factorV = 2 * B / (V * V * V)
                gV += factorV * DP * self.pOld
            factorB = -1 / (V * V)
            gB += factorB * DP * self.pOld - t1[0] * t2[0] - t1[1] * t2[1] - t1[2] * t2[2]

    def adjoint_born_accumulation_VB(self, gV, gB, P, DP):
        self._acc_vb(gV, gB, P, DP)

    def adjoint_born_accumulation_B(self, gB, P, DP):
        self._acc_vb(None, gB, P, DP)

    def adjoint_born_accumulation_wavefieldsep(self, gV, DP, isFWI):
        """src/prop3DAcoIsoDenQ_DEO2_FDTD.h:673-765"""
        with np.errstate(all="ignore"):
            factor = 2 * self.b / (self.v * self.v * self.v)
        gV += (factor * wavefield_separation(DP, self.pOld, isFWI)).astype(F32)


class OracleVTI(_OracleProp):
    """Prop3DAcoVTIDenQ_DEO2_FDTD (src/prop3DAcoVTIDenQ_DEO2_FDTD.h:16-1897)"""

    MODEL = ("v", "eps", "eta", "b", "f")
    FIELDS = ("pCur", "pOld", "mCur", "mOld", "pSpace", "mSpace")
    SPACE = ("pSpace", "mSpace")

    def _tmp(self):
        return np.empty(6 * self.v.size, F32)

    def time_step(self, linear=False):
        """src/prop3DAcoVTIDenQ_DEO2_FDTD.h:326-376 (linear: timeStepLinear)"""
        g = self.g
        lib().wo_vti_step(*g.head(), g.dt, _p(self.v), _p(self.eps), _p(self.eta), _p(self.b), _p(self.f),
                          _p(self.q), _p(self.pCur), _p(self.mCur),
                          None if linear else _p(self.pSpace), None if linear else _p(self.mSpace),
                          _p(self.pOld), _p(self.mOld), _p(self._tmp()), *g.tail())
        self.pCur, self.pOld = self.pOld, self.pCur
        self.mCur, self.mOld = self.mOld, self.mCur

    def time_step_linear(self):
        self.time_step(linear=True)

    def forward_born_injection_V(self, dV, DP, DM):
        """src/prop3DAcoVTIDenQ_DEO2_FDTD.h:420-452"""
        with np.errstate(all="ignore"):
            dt = self.g.dt
            factor = 2 * dt * dt * dV / self.v
            self.pCur += factor * DP
            self.mCur += factor * DM

    def _rplus(self, P):
        return dplus(self.g, P, P, P)

    def _rminus(self, a, b, c):
        return dminus(self.g, a, b, c)

    def forward_born_injection_VEA(self, dV, dE, dA, P, M, DP, DM):
        """src/prop3DAcoVTIDenQ_DEO2_FDTD.h:457-551 (TTI: :460-554 with rotated D+/D-)"""
        tP1 = self._rplus(P)
        tM1 = self._rplus(M)
        with np.errstate(all="ignore"):
            A, B, F = self.eta, self.b, self.f
            coef = dA * B * F * (1 - 2 * A * A) / np.sqrt(1 - A * A)
            tP2 = [(2 * B * dE) * tP1[0], (2 * B * dE) * tP1[1],
                   (-2 * B * F * A * dA) * tP1[2] + coef * tM1[2]]
            zero = self.g.zeros()
            tM2 = [zero, zero, (2 * B * F * A * dA) * tM1[2] + coef * tP1[2]]
        tP1 = self._rminus(*tP2)
        tM1 = self._rminus(*tM2)
        with np.errstate(all="ignore"):
            V = self.v
            dt2v2OverB = self._dt2v2OverB()
            factor = 2 * B * dV / (V * V * V)
            self.pCur += dt2v2OverB * (factor * DP + tP1[0] + tP1[1] + tP1[2])
            self.mCur += dt2v2OverB * (factor * DM + tM1[0] + tM1[1] + tM1[2])

    def adjoint_born_accumulation_V(self, gV, DP, DM):
        """src/prop3DAcoVTIDenQ_DEO2_FDTD.h:564-589"""
        with np.errstate(all="ignore"):
            factor = 2 * self.b / (self.v * self.v * self.v)
            gV += factor * (DP * self.pOld + DM * self.mOld)

    def adjoint_born_accumulation_VEA(self, gV, gE, gA, P, M, DP, DM):
        """src/prop3DAcoVTIDenQ_DEO2_FDTD.h:730-787 (TTI: :733-790)"""
        tP1 = self._rplus(P)
        tM1 = self._rplus(M)
        tP2 = self._rplus(self.pOld)
        tM2 = self._rplus(self.mOld)
        with np.errstate(all="ignore"):
            V, A, B, F = self.v, self.eta, self.b, self.f
            factor = 2 * B / (V * V * V)
            gV += factor * (DP * self.pOld + DM * self.mOld)
            gE += (-2 * B * tP1[0] * tP2[0] - 2 * B * tP1[1] * tP2[1])
            c = B * F * (1 - 2 * A * A) / np.sqrt(1 - A * A)
            partP = 2 * B * F * A * tP1[2] - c * tM1[2]
            partM = 2 * B * F * A * tM1[2] + c * tP1[2]
            gA += (partP * tP2[2] - partM * tM2[2])

    def adjoint_born_accumulation_wavefieldsep_V(self, gV, DP, DM, isFWI):
        """src/prop3DAcoVTIDenQ_DEO2_FDTD.h:606-725"""
        with np.errstate(all="ignore"):
            factor = 2 * self.b / (self.v * self.v * self.v)
        s = wavefield_separation(DP, self.pOld, isFWI) + wavefield_separation(DM, self.mOld, isFWI)
        gV += (factor * s).astype(F32)


class OracleTTI(OracleVTI):
    """Prop3DAcoTTIDenQ_DEO2_FDTD (src/prop3DAcoTTIDenQ_DEO2_FDTD.h:14-3659)"""

    MODEL = ("v", "eps", "eta", "b", "f", "sinTheta", "cosTheta", "sinPhi", "cosPhi")

    def time_step(self, linear=False):
        """src/prop3DAcoTTIDenQ_DEO2_FDTD.h:359-379"""
        assert not linear, "the TTI propagator has no linear step"
        g = self.g
        lib().wo_tti_step(*g.head(), g.dt, _p(self.v), _p(self.eps), _p(self.eta), _p(self.b), _p(self.f),
                          _p(self.q), _p(self.sinTheta), _p(self.cosTheta), _p(self.sinPhi), _p(self.cosPhi),
                          _p(self.pCur), _p(self.mCur), _p(self.pSpace), _p(self.mSpace),
                          _p(self.pOld), _p(self.mOld), _p(self._tmp()), *g.tail())
        self.pCur, self.pOld = self.pOld, self.pCur
        self.mCur, self.mOld = self.mOld, self.mCur

    def _angles(self):
        return self.sinTheta, self.cosTheta, self.sinPhi, self.cosPhi

    def _rplus(self, P):
        return tti_dplus_rot(self.g, P, P, P, *self._angles())

    def _rminus(self, a, b, c):
        return tti_dminus_rot(self.g, a, b, c, *self._angles())


# --------------------------------------------------------------------------------------
# wavefield separation (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:673-765)
# --------------------------------------------------------------------------------------

def wavefield_separation(nlf, adj, isFWI):
    """Per (x,y) column: zero-pad to 2nz, transform with FFTW's sign +1, zero half of the
    spectrum (nonlinear: bins nz+1..2nz-1; adjoint: 1..nz-1 for FWI, nz+1..2nz-1 for RTM),
    transform back with sign -1, and return Re(S*R) on the first nz samples (float64)."""
    nx, ny, nz = nlf.shape
    nfft = 2 * nz
    scale = 1.0 / nfft
    a = np.zeros((nx, ny, nfft), np.float64)
    b = np.zeros((nx, ny, nfft), np.float64)
    a[..., :nz] = scale * nlf.astype(np.float64)
    b[..., :nz] = scale * adj.astype(np.float64)
    # FFTW sign +1 is numpy's ifft times n
    A = np.fft.ifft(a, axis=-1) * nfft
    Bs = np.fft.ifft(b, axis=-1) * nfft
    A[..., nz + 1:] = 0
    kadj = 0 if isFWI else nz
    Bs[..., kadj + 1:kadj + nz] = 0
    # FFTW sign -1 is numpy's (unnormalised) fft
    a2 = np.fft.fft(A, axis=-1)[..., :nz]
    b2 = np.fft.fft(Bs, axis=-1)[..., :nz]
    return np.real(a2 * b2)


# --------------------------------------------------------------------------------------
# injection / extraction / interpolation
# --------------------------------------------------------------------------------------

SOURCE_POINT32 = np.dtype([("iu", np.int64), ("ir", np.int64), ("c", np.float32)], align=True)


def inject(field, data, it, points):
    """src/spacetime.cc:144-162; data is (nr, nt) C-order == Julia (nt, nr)."""
    pts = np.ascontiguousarray(points, dtype=SOURCE_POINT32)
    nt = data.shape[1]
    lib().wo_inject(_p(field.reshape(-1)), _p(np.ascontiguousarray(data, F32).reshape(-1)), it, nt,
                    pts.ctypes.data, len(pts))


def extract(data, field, it, points):
    """src/spacetime.cc:184-204 (accumulates)."""
    pts = np.ascontiguousarray(points, dtype=SOURCE_POINT32)
    nt = data.shape[1]
    lib().wo_extract(_p(data.reshape(-1)), _p(np.ascontiguousarray(field, F32).reshape(-1)),
                     pts.ctypes.data, len(pts), it, nt)


def interp_forward(h, m):
    """fine -> coarse, src/spacetime.cc:72-127; h: (nh, 8); m: (nslow, nm)"""
    h = np.ascontiguousarray(h, F32)
    m = np.ascontiguousarray(m, F32)
    nh = h.shape[0]
    nslow, nm = m.shape
    nd = -(-nm // nh)
    d = np.zeros((nslow, nd), F32)
    lib().wo_interp_forward(_p(h), _p(d), _p(m), nh, nd, nm, nslow)
    return d


def interp_adjoint(h, d, nm):
    """coarse -> fine, src/spacetime.cc:4-63"""
    h = np.ascontiguousarray(h, F32)
    d = np.ascontiguousarray(d, F32)
    nh = h.shape[0]
    nslow, nd = d.shape
    m = np.zeros((nslow, nm), F32)
    lib().wo_interp_adjoint(_p(h), _p(m), _p(d), nh, nm, nd, nslow)
    return m
