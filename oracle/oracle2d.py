"""CPU oracle for the 2D Iso propagator -- TEST INFRASTRUCTURE ONLY.

Wraps the 2D functions of ``oracle/lib/libwavefd_oracle.so`` (``wo2_*``, a C restatement of
src/propagatorStaticFunctions.h:10-311, src/prop2DAcoIsoDenQ_DEO2_FDTD.h:674-1045 and the 2D Q
profile of src/absorb.h:23-75) and composes the Born sources and imaging conditions
(src/prop2DAcoIsoDenQ_DEO2_FDTD.h:249-582) in float32 numpy in the reference's expression order.
Arrays are C-ordered (nx, nz) float32, i.e. k = kx*nz + kz (Julia's (nz, nx) column-major view).
The product package never imports it.

Pins (tests/test_oracle2d.py): D+/D- exact on polynomials of degree <= 7 along x and z and
discretely adjoint (<D+u, w> = -<u, D-w>) with and without the free-surface mirrors; the Iso
step equals the 3D oracle's on an (nx, 9, nz) grid whose y-derivatives vanish (a field
constant in y) in the rows the 3D interior covers; energy conservation of the Q = 0 leapfrog;
the Q profile against its closed form.
"""
import ctypes

import numpy as np

from oracle import oracle as _o

F32 = np.float32
_L, _F = ctypes.c_long, ctypes.c_float
_fp = ctypes.POINTER(ctypes.c_float)


def lib():
    l = _o.lib()
    if not getattr(l, "_wo2_declared", False):
        l.wo2_setup_q.argtypes = [_L, _L, _L, _L, _F, _F, _F, _F, _fp]
        l.wo2_setup_q.restype = ctypes.c_int
        d = [_L, _L, _L, _F, _F, _fp, _fp, _fp, _fp]
        l.wo2_dplus.argtypes = d
        l.wo2_dminus.argtypes = d
        l.wo2_iso_sandwich.argtypes = [_L, _L, _L, _F, _F, _fp, _fp, _fp, _fp]
        l.wo2_iso_update.argtypes = [_L, _L, _L, _F, _F, _F] + [_fp] * 8
        l._wo2_declared = True
    return l


def _p(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


class Grid2:
    def __init__(self, nx, nz, dx, dz, fs):
        self.nx, self.nz, self.fs = nx, nz, 1 if fs else 0
        # 1/d in double, stored to float (src/prop2DAcoIsoDenQ_DEO2_FDTD.h:61-62)
        self.idx, self.idz = F32(1.0 / F32(dx)), F32(1.0 / F32(dz))

    def head(self):
        return (self.fs, self.nx, self.nz, self.idx, self.idz)


def dplus(g, inX, inZ):
    ox, oz = np.empty_like(inX), np.empty_like(inZ)
    lib().wo2_dplus(*g.head(), _p(inX), _p(inZ), _p(ox), _p(oz))
    return ox, oz


def dminus(g, inX, inZ):
    ox, oz = np.empty_like(inX), np.empty_like(inZ)
    lib().wo2_dminus(*g.head(), _p(inX), _p(inZ), _p(ox), _p(oz))
    return ox, oz


def setup_q(fs, nx, nz, nsponge, dt, freqQ, qMin, qInterior):
    q = np.zeros((nx, nz), F32)
    if lib().wo2_setup_q(1 if fs else 0, nx, nz, nsponge, F32(dt), F32(freqQ), F32(qMin), F32(qInterior), _p(q)):
        raise ValueError("freqQ too small")
    return q


class Oracle2DIso:
    """Prop2DAcoIsoDenQ_DEO2_FDTD (src/prop2DAcoIsoDenQ_DEO2_FDTD.h:16-1048)"""

    def __init__(self, nx, nz, dx, dz, dt, fs=True, nsponge=60, freqQ=5.0, qMin=0.1, qInterior=100.0):
        self.g = Grid2(nx, nz, dx, dz, fs)
        self.dt = F32(dt)
        for name in ("v", "b", "pCur", "pOld", "pSpace"):
            setattr(self, name, np.zeros((nx, nz), F32))
        self.q = setup_q(fs, nx, nz, nsponge, dt, freqQ, qMin, qInterior)

    def time_step(self):
        """src/prop2DAcoIsoDenQ_DEO2_FDTD.h:200-216"""
        g = self.g
        tX, tZ = np.empty_like(self.v), np.empty_like(self.v)
        lib().wo2_iso_sandwich(*g.head(), _p(self.pCur), _p(self.b), _p(tX), _p(tZ))
        lib().wo2_iso_update(*g.head(), self.dt, _p(tX), _p(tZ), _p(self.v), _p(self.b), _p(self.q),
                             _p(self.pCur), _p(self.pSpace), _p(self.pOld))
        self.pCur, self.pOld = self.pOld, self.pCur

    def scale_spatial_derivatives(self):
        """:222-239"""
        v2OverB = self.v * self.v / self.b
        self.pSpace *= v2OverB

    # Born sources :249-432
    def forward_born_injection_V(self, dV, DP):
        with np.errstate(all="ignore"):
            factor = 2 * self.dt * self.dt * dV / self.v
            self.pCur += factor * DP

    def _lb(self, dB, P):
        g = self.g
        tx, tz = dplus(g, P, P)
        return dminus(g, dB * tx, dB * tz)

    def forward_born_injection_VB(self, dV, dB, P, DP):
        lx, lz = self._lb(dB, P)
        with np.errstate(all="ignore"):
            V, B = self.v, self.b
            V2 = V * V
            dt2v2OverB = self.dt * self.dt * V2 / B
            self.pCur += dt2v2OverB * (DP / V2 * (2 * B * dV / V - dB) + lx + lz)

    def forward_born_injection_B(self, dB, P, DP):
        lx, lz = self._lb(dB, P)
        with np.errstate(all="ignore"):
            V, B = self.v, self.b
            V2 = V * V
            dt2v2OverB = self.dt * self.dt * V2 / B
            factorB = -dB / V2
            self.pCur += dt2v2OverB * (DP * factorB + lx + lz)

    # imaging :444-582
    def adjoint_born_accumulation_V(self, gV, DP):
        with np.errstate(all="ignore"):
            factorV = +2 * self.b / (self.v * self.v * self.v)
            gV += factorV * DP * self.pOld

    def _acc(self, gV, gB, P, DP):
        g = self.g
        px, pz = dplus(g, P, P)
        qx, qz = dplus(g, self.pOld, self.pOld)
        with np.errstate(all="ignore"):
            V, B = self.v, self.b
            if gV is not None:
                factorV = +2 * B / (V * V * V)
                gV += factorV * DP * self.pOld
            factorB = -1 / (V * V)
            gB += factorB * DP * self.pOld - px * qx - pz * qz

    def adjoint_born_accumulation_VB(self, gV, gB, P, DP):
        self._acc(gV, gB, P, DP)

    def adjoint_born_accumulation_B(self, gB, P, DP):
        self._acc(None, gB, P, DP)
