"""CPU oracle for the 2D Iso propagator -- TEST INFRASTRUCTURE ONLY.

Wraps the 2D functions of ``oracle/lib/libwavefd_oracle.so`` (``wo2_*``, a C restatement of
src/propagatorStaticFunctions.h:10-311, src/prop2DAcoIsoDenQ_DEO2_FDTD.h:674-1045 and the 2D Q
profile of src/absorb.h:23-75) and composes the Born sources and imaging conditions
(src/prop2DAcoIsoDenQ_DEO2_FDTD.h:249-582) in float32 numpy in the reference's expression order.
Arrays are C-ordered (nx, nz) float32, i.e. k = kx*nz + kz (Julia's (nz, nx) column-major view).
The product package never imports it.

The VTI and TTI classes restate src/prop2DAcoVTIDenQ_DEO2_FDTD.h:16-1478 and
src/prop2DAcoTTIDenQ_DEO2_FDTD.h:14-1717 the same way (``wo2_vti_*``, ``wo2_tti_*``).

Pins (tests/test_oracle2d.py): D+/D- exact on polynomials of degree <= 7 along x and z and
discretely adjoint (<D+u, w> = -<u, D-w>) with and without the free-surface mirrors; the Iso
step equals the 3D oracle's on an (nx, 9, nz) grid whose y-derivatives vanish (a field
constant in y) in the rows the 3D interior covers; the Q profile against its closed form;
VTI: eps = eta = 0 gives the Iso step on P bit for bit, and the step equals the 3D VTI
oracle's on y-constant fields (bit for bit at eta = 0, where the 2D and 3D texts' different
grouping of the coupling term cannot show, and to 1e-6 otherwise); TTI: theta = 0 gives the
VTI step bit for bit, theta = 90 degrees the VTI step on the transposed grid bit for bit, and
the rotated D+ and D- are negative adjoints; every Born source is the finite-difference
derivative of the step and every imaging condition its adjoint.
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
        l.wo2_vti_sandwich.argtypes = [_L, _L, _L, _F, _F] + [_fp] * 10
        l.wo2_vti_update.argtypes = [_L, _L, _L, _F, _F, _F] + [_fp] * 13
        l.wo2_tti_dplus.argtypes = [_L, _L, _L, _F, _F] + [_fp] * 6
        l.wo2_tti_dminus.argtypes = [_L, _L, _L, _F, _F] + [_fp] * 6
        l.wo2_tti_sandwich.argtypes = [_L, _L, _L, _F, _F] + [_fp] * 12
        l.wo2_tti_update.argtypes = [_L, _L, _L, _F, _F, _F] + [_fp] * 15
        l._wo2_declared = True
    return l


def _p(a):
    if a is None:
        return None
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


def setup_q(fs, nx, nz, nsponge, dt, freqQ, qMin, qInterior, out=None):
    q = np.zeros((nx, nz), F32) if out is None else out
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


class Oracle2DVTI(Oracle2DIso):
    """Prop2DAcoVTIDenQ_DEO2_FDTD (src/prop2DAcoVTIDenQ_DEO2_FDTD.h:16-1478)"""

    def __init__(self, nx, nz, dx, dz, dt, **kw):
        super().__init__(nx, nz, dx, dz, dt, **kw)
        for name in ("eps", "eta", "f", "mCur", "mOld", "mSpace"):
            setattr(self, name, np.zeros((nx, nz), F32))

    def _z(self):
        return np.empty_like(self.v)

    def _sandwich(self):
        g = self.g
        t = [self._z() for _ in range(4)]
        lib().wo2_vti_sandwich(*g.head(), _p(self.pCur), _p(self.mCur), _p(self.eps), _p(self.eta), _p(self.f),
                               _p(self.b), *[_p(a) for a in t])
        return t

    def _update(self, t, linear):
        g = self.g
        lib().wo2_vti_update(*g.head(), self.dt, *[_p(a) for a in t], _p(self.v), _p(self.b), _p(self.q),
                             _p(self.pCur), _p(self.mCur), None if linear else _p(self.pSpace),
                             None if linear else _p(self.mSpace), _p(self.pOld), _p(self.mOld))

    def time_step(self, linear=False):
        """:263-314 (linear: timeStepLinear)"""
        self._update(self._sandwich(), linear)
        self.pCur, self.pOld = self.pOld, self.pCur
        self.mCur, self.mOld = self.mOld, self.mCur

    def time_step_linear(self):
        self.time_step(linear=True)

    def scale_spatial_derivatives(self):
        """:320-338"""
        v2OverB = self.v * self.v / self.b
        self.pSpace *= v2OverB
        self.mSpace *= v2OverB

    def forward_born_injection_V(self, dV, DP, DM):
        """:351-375"""
        with np.errstate(all="ignore"):
            factor = 2 * self.dt * self.dt * dV / self.v
            self.pCur += factor * DP
            self.mCur += factor * DM

    def _dplus(self, a):
        return dplus(self.g, a, a)

    def _dminus(self, x, z):
        return dminus(self.g, x, z)

    def forward_born_injection_VEA(self, dV, dE, dA, P, M, DP, DM):
        """:383-462"""
        tPx1, tPz1 = self._dplus(P)
        tMx1, tMz1 = self._dplus(M)
        with np.errstate(all="ignore"):
            A, B, F = self.eta, self.b, self.f
            tPx2 = (2 * B * dE) * tPx1
            tPz2 = (-2 * B * F * A * dA) * tPz1 + (dA * B * F * (1 - 2 * A * A) / np.sqrt(1 - A * A)) * tMz1
            tMx2 = np.zeros_like(tPx1)
            tMz2 = (2 * B * F * A * dA) * tMz1 + (dA * B * F * (1 - 2 * A * A) / np.sqrt(1 - A * A)) * tPz1
        tPx1, tPz1 = self._dminus(tPx2, tPz2)
        tMx1, tMz1 = self._dminus(tMx2, tMz2)
        with np.errstate(all="ignore"):
            V = self.v
            dt2v2OverB = self.dt * self.dt * V * V / B
            factor = 2 * B * dV / (V * V * V)
            self.pCur += dt2v2OverB * (factor * DP + tPx1 + tPz1)
            self.mCur += dt2v2OverB * (factor * DM + tMx1 + tMz1)

    def adjoint_born_accumulation_V(self, gV, DP, DM):
        """:473-492"""
        with np.errstate(all="ignore"):
            factor = 2 * self.b / (self.v * self.v * self.v)
            gV += factor * (DP * self.pOld + DM * self.mOld)

    def adjoint_born_accumulation_VEA(self, gV, gE, gA, P, M, DP, DM):
        """:629-686"""
        tPx1, tPz1 = self._dplus(P)
        tMx1, tMz1 = self._dplus(M)
        tPx2, tPz2 = self._dplus(self.pOld)
        tMx2, tMz2 = self._dplus(self.mOld)
        with np.errstate(all="ignore"):
            V, A, B, F = self.v, self.eta, self.b, self.f
            factor = 2 * B / (V * V * V)
            gV += factor * DP * self.pOld + factor * DM * self.mOld
            gE += -2 * B * tPx1 * tPx2
            c = B * F * (1 - 2 * A * A) / np.sqrt(1 - A * A)
            partP = 2 * B * F * A * tPz1 - c * tMz1
            partM = 2 * B * F * A * tMz1 + c * tPz1
            gA += partP * tPz2 - partM * tMz2


class Oracle2DTTI(Oracle2DVTI):
    """Prop2DAcoTTIDenQ_DEO2_FDTD (src/prop2DAcoTTIDenQ_DEO2_FDTD.h:14-1717)"""

    def __init__(self, nx, nz, dx, dz, dt, **kw):
        super().__init__(nx, nz, dx, dz, dt, **kw)
        self.sinTheta = np.zeros((nx, nz), F32)
        self.cosTheta = np.zeros((nx, nz), F32)

    def _sandwich(self):
        g = self.g
        t = [self._z() for _ in range(4)]
        lib().wo2_tti_sandwich(*g.head(), _p(self.pCur), _p(self.mCur), _p(self.sinTheta), _p(self.cosTheta),
                               _p(self.eps), _p(self.eta), _p(self.f), _p(self.b), *[_p(a) for a in t])
        return t

    def _update(self, t, linear):
        assert not linear, "the 2D TTI propagator has no linear step"
        g = self.g
        lib().wo2_tti_update(*g.head(), self.dt, *[_p(a) for a in t], _p(self.sinTheta), _p(self.cosTheta),
                             _p(self.v), _p(self.b), _p(self.q), _p(self.pCur), _p(self.mCur), _p(self.pSpace),
                             _p(self.mSpace), _p(self.pOld), _p(self.mOld))

    def _dplus(self, a):
        return tti_dplus(self.g, a, a, self.sinTheta, self.cosTheta)

    def _dminus(self, x, z):
        return tti_dminus(self.g, x, z, self.sinTheta, self.cosTheta)


def tti_dplus(g, g1, g3, sT, cT):
    """rotated D+ (src/prop2DAcoTTIDenQ_DEO2_FDTD.h:1366-1528)"""
    o1, o3 = np.empty_like(g1), np.empty_like(g3)
    lib().wo2_tti_dplus(*g.head(), _p(g1), _p(g3), _p(sT), _p(cT), _p(o1), _p(o3))
    return o1, o3


def tti_dminus(g, g1, g3, sT, cT):
    """rotated D- (src/prop2DAcoTTIDenQ_DEO2_FDTD.h:1531-1716)"""
    o1, o3 = np.empty_like(g1), np.empty_like(g3)
    lib().wo2_tti_dminus(*g.head(), _p(g1), _p(g3), _p(sT), _p(cT), _p(o1), _p(o3))
    return o1, o3
