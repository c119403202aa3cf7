"""Source injection, receiver extraction and time interpolation (host API).

Mirrors src/spacetime.jl: ``hickscoeffs`` / ``linearcoeffs`` build ``SourcePoint32`` taps
(1-based linear grid index ``iu``, 1-based trace ``ir``, coefficient ``c``) with the
reference's integer index rules, ``source_blocking`` / ``receiver_blocking`` partition them
so that no key is split across partitions, ``injectdata`` / ``extractdata`` call
libspacetime's ``injectdata_float`` / ``extractdata_float``, and ``interpfilters`` /
``interpforward`` / ``interpadjoint`` build and apply the 8-tap sinc filters.

Arrays follow Julia's memory layout: a wavefield is Julia ``(nz, ny, nx)`` column-major,
i.e. any float32 buffer of nx*ny*nz values with z fastest; trace data Julia ``(nt, nr)``
column-major is numpy ``(nr, nt)`` C-order.
"""
import ctypes

import numpy as np

from . import _lib

SourcePoint32 = np.dtype([("iu", np.int64), ("ir", np.int64), ("c", np.float32)], align=True)
SourcePoint64 = np.dtype([("iu", np.int64), ("ir", np.int64), ("c", np.float64)], align=True)

assert SourcePoint32.itemsize == 24 and SourcePoint64.itemsize == 24


# ---- coefficients ------------------------------------------------------------------------

def _f32(x):
    return np.float32(x)


def kaiser(b, r, x):
    """src/spacetime.jl:678: I0(b sqrt(1-(x/r)^2)) / I0(b) for |x| <= r, else 0; the argument
    is formed in float32 like the reference's Float32 path, the Bessel ratio in float64
    (SpecialFunctions evaluates besseli(0, ::Float32) through Float64) and rounded."""
    x = np.asarray(x, np.float32)
    b = np.float32(b)
    inside = np.abs(x) <= r
    t = (np.float32(1) - (x / np.float32(r)) ** 2).astype(np.float32)
    arg = (b * np.sqrt(np.maximum(t, np.float32(0)))).astype(np.float32)
    num = np.i0(arg.astype(np.float64)).astype(np.float32)
    den = np.float32(np.i0(np.float64(b)))
    return np.where(inside, num / den, np.float32(0)).astype(np.float32)


def sinc32(x):
    """sin(pi x)/(pi x) in float32 (1 at 0)."""
    x = np.asarray(x, np.float32)
    s = np.sin(np.pi * x.astype(np.float64)).astype(np.float32)
    d = (np.float32(np.pi) * x).astype(np.float32)
    with np.errstate(all="ignore"):
        out = (s / d).astype(np.float32)
    return np.where(x == 0, np.float32(1), out).astype(np.float32)


def _hicks_1d(b, r, alpha):
    """the 2r taps at offsets -r+1..r of one offset alpha (shape (2r,)) or of an array of them
    (shape alpha.shape + (2r,))"""
    a = np.arange(1, 2 * r + 1) - r  # i - r for i = 1..2r  -> -3..4
    z = (a.astype(np.float32) - np.asarray(alpha, np.float32)[..., None]).astype(np.float32)
    if np.ndim(alpha) == 0:
        z = z.reshape(-1)
    return (kaiser(b, r, z) * sinc32(z)).astype(np.float32)


def hickscoeffs(dz, dy, dx, z0, y0, x0, nz, ny, nx, z, y, x, fs_index=1, zstagger=0.0, ystagger=0.0, xstagger=0.0):
    """3D Kaiser-windowed sinc taps, src/spacetime.jl:867-925.  dz,dy,dx are the element type
    (float32 here); z,y,x float64 positions.  Returns a SourcePoint32 array.  All points at once:
    per point the 8 x 8 x 8 taps in the reference's loop order (iz fastest, then iy, ix), with
    the reference's float32 products and free-surface image terms."""
    z, y, x = (np.atleast_1d(np.asarray(v, np.float64)) for v in (z, y, x))
    hb, hr = np.float32(6.42), 4
    hl = 2 * hr
    dzd, dyd, dxd = float(np.float32(dz)), float(np.float32(dy)), float(np.float32(dx))
    npt = len(z)
    iz_mid = np.floor((z - z0) / dzd).astype(np.int64) + 1
    iy_mid = np.floor((y - y0) / dyd).astype(np.int64) + 1
    ix_mid = np.floor((x - x0) / dxd).astype(np.int64) + 1
    az = ((z - (z0 + (iz_mid - 1 + zstagger) * dzd)) / dzd).astype(np.float32)
    ay = ((y - (y0 + (iy_mid - 1 + ystagger) * dyd)) / dyd).astype(np.float32)
    ax = ((x - (x0 + (ix_mid - 1 + xstagger) * dxd)) / dxd).astype(np.float32)
    fz, fy, fx = _hicks_1d(hb, hr, az), _hicks_1d(hb, hr, ay), _hicks_1d(hb, hr, ax)  # (npt, 8)
    # d[p,i,j,k] = fz[p,i]*fy[p,j]*fx[p,k] in float32
    ci = (fz[:, :, None, None] * fy[:, None, :, None]).astype(np.float32) * fx[:, None, None, :]
    ci = ci.astype(np.float32)
    # reflect source terms above the free surface (src/spacetime.jl:905-917)
    iz0 = iz_mid - hr + 1
    m = iz0 == fs_index - 3
    ci[m, 4] -= ci[m, 2]
    ci[m, 5] -= ci[m, 1]
    ci[m, 6] -= ci[m, 0]
    m = iz0 == fs_index - 2
    ci[m, 3] -= ci[m, 1]
    ci[m, 4] -= ci[m, 0]
    m = iz0 == fs_index - 1
    ci[m, 2] -= ci[m, 0]
    off = np.arange(1, hl + 1) - hr
    shape = (npt, hl, hl, hl)
    IZ = np.broadcast_to((iz_mid[:, None] + off[None, :])[:, :, None, None], shape)
    IY = np.broadcast_to((iy_mid[:, None] + off[None, :])[:, None, :, None], shape)
    IX = np.broadcast_to((ix_mid[:, None] + off[None, :])[:, None, None, :], shape)

    def order(a):  # per point, iz fastest
        return np.ascontiguousarray(a.transpose(0, 3, 2, 1)).ravel()

    ir = np.repeat(np.arange(1, npt + 1), hl ** 3)
    return _finish(order(IZ), order(IY), order(IX), ir, order(ci), nz, ny, nx)


def linearcoeffs(dz, dy, dx, z0, y0, x0, nz, ny, nx, z, y, x, fs_index=1):
    """3D trilinear taps, src/spacetime.jl:967-1013, all points at once: per point the 8 taps in
    the reference's loop order (iz fastest, then iy, ix), weights wx*wy*wz in float64 rounded to
    float32, the free-surface row zeroed."""
    z, y, x = (np.atleast_1d(np.asarray(v, np.float64)) for v in (z, y, x))
    dzd, dyd, dxd = float(np.float32(dz)), float(np.float32(dy)), float(np.float32(dx))
    kx = np.floor((x - x0) / dxd).astype(np.int64) + 1
    ky = np.floor((y - y0) / dyd).astype(np.int64) + 1
    kz = np.floor((z - z0) / dzd).astype(np.int64) + 1
    rx = (x - (x0 + dxd * (kx - 1))) / dxd
    ry = (y - (y0 + dyd * (ky - 1))) / dyd
    rz = (z - (z0 + dzd * (kz - 1))) / dzd
    npt = len(z)
    iz, iy, ix = (np.empty((npt, 8), np.int64) for _ in range(3))
    c = np.empty((npt, 8), np.float32)
    fsrow = kz == fs_index  # the free surface row gets no weight (src/spacetime.jl:1000-1005)
    for cc in range(2):
        wx = rx if cc else 1.0 - rx
        for bb in range(2):
            wy = ry if bb else 1.0 - ry
            for a in range(2):
                wz = rz if a else 1.0 - rz
                j = a + 2 * bb + 4 * cc
                iz[:, j], iy[:, j], ix[:, j] = kz + a, ky + bb, kx + cc
                cj = (wx * wy * wz).astype(np.float32)
                if a == 0:
                    cj[fsrow] = 0
                c[:, j] = cj
    ir = np.repeat(np.arange(1, npt + 1), 8)
    return _finish(iz.ravel(), iy.ravel(), ix.ravel(), ir, c.ravel(), nz, ny, nx)


def _finish(iz, iy, ix, ir, c, nz, ny, nx):
    ok = (iz >= 1) & (iz <= nz) & (iy >= 1) & (iy <= ny) & (ix >= 1) & (ix <= nx)
    iz, iy, ix, ir, c = iz[ok], iy[ok], ix[ok], ir[ok], c[ok]
    keep = np.abs(c) >= np.finfo(np.float32).eps
    iz, iy, ix, ir, c = iz[keep], iy[keep], ix[keep], ir[keep], c[keep]
    pts = np.zeros(len(iz), SourcePoint32)
    # LinearIndices((nz,ny,nx))[iz,iy,ix], 1-based
    pts["iu"] = iz + (iy - 1) * nz + (ix - 1) * nz * ny
    pts["ir"] = ir
    pts["c"] = c
    return pts


# ---- blocking (src/spacetime.jl:1049-1129) ---------------------------------------------

def blocking(points, key, nthreads):
    points = np.asarray(points)
    if len(points) == 0 or nthreads < 1:
        raise ValueError("max(nthreads, length(points)) must be greater than zero")
    order = np.argsort(points[key], kind="stable")
    pts = points[order]
    npart = min(nthreads, len(pts))
    nominal, rem = divmod(len(pts), npart)
    if rem > 0:
        nominal += 1
    parts = [[] for _ in range(npart)]
    ip, counter = 0, 1
    for i in range(len(pts)):
        if counter <= nominal or (i > 0 and pts[key][i] == pts[key][i - 1]):
            parts[ip].append(i)
            counter += 1
        else:
            ip += 1
            if ip >= len(parts):
                parts.append([])
            parts[ip].append(i)
            counter = 1
    return [pts[np.array(p, dtype=np.int64)] for p in parts if p]


def source_blocking(points, nthreads=8):
    return blocking(points, "iu", nthreads)


def receiver_blocking(points, nthreads=8):
    return blocking(points, "ir", nthreads)


def _partition_arrays(partitions):
    keep = [np.ascontiguousarray(p, dtype=SourcePoint32) for p in partitions]
    ptrs = (ctypes.c_void_p * len(keep))(*[k.ctypes.data for k in keep])
    npts = np.array([len(k) for k in keep], np.int64)
    keep.append(ptrs)
    keep.append(npts)
    return keep, ctypes.addressof(ptrs), len(partitions), npts.ctypes.data


def _addr(a):
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return int(a.data_ptr())
    assert a.dtype == np.float32 and (a.flags.c_contiguous or a.flags.f_contiguous)
    return a.ctypes.data


def injectdata(field, partitions, data, it, nthreads=8):
    """field += c * data[it] at the taps, src/spacetime.jl:1157-1162.  data: (nr, nt) C-order
    (Julia (nt, nr)); it is 1-based.  field may be a host array or a device address."""
    keep, pp, nparts, npts = _partition_arrays(partitions)
    nt = data.shape[-1]
    _lib.load("spacetime").injectdata_float(_addr(field), _addr(data), it, nt, pp, nparts, npts, nthreads)


def extractdata(data, field, it, partitions, nthreads=8):
    """data[it] += c * field at the taps (accumulates), src/spacetime.jl:1182-1187."""
    keep, pp, nparts, npts = _partition_arrays(partitions)
    nt = data.shape[-1]
    _lib.load("spacetime").extractdata_float(_addr(data), _addr(field), pp, nparts, npts, it, nt, nthreads)


# ---- time interpolation (src/spacetime.jl:333-660) --------------------------------------

def interpfilters(dtmod, dtrec, mode=0):
    """8-tap sinc filters mapping dtmod <-> dtrec, src/spacetime.jl:333-364.  Returns (n, 8)
    float32 rows h[i] = A \\ c_i (solved in float64, then rounded), divided by n for mode 0."""
    assert dtmod <= dtrec
    n = int(round(dtrec / dtmod))
    if abs(n * dtmod - dtrec) / dtrec > 1e-6:
        raise ValueError("abs(n*dtmod-dtrec)/dtrec > 1e-6")
    d = dtmod / dtrec * np.arange(n)
    fmax = 0.066 + 0.265 * np.log(8.0)
    j = np.arange(1, 9)
    a = np.sinc((j - 1) * fmax)
    A = np.array([[a[abs(r - c)] for c in range(8)] for r in range(8)])
    h = []
    for i in range(n):
        c = np.sinc((4 - j + d[i]) * fmax)
        hi = np.linalg.solve(A, c)
        if mode == 0:
            hi = hi / n
        h.append(hi)
    return np.array(h, np.float32)


def _hptrs(h):
    rows = [np.ascontiguousarray(r) for r in h]
    arr = (ctypes.c_void_p * len(rows))(*[r.ctypes.data for r in rows])
    return rows, arr


def interpforward(h, d, m):
    """fine m (nslow, nm) -> coarse d (nslow, nd), src/spacetime.jl:606-618."""
    if m.shape[-1] == d.shape[-1]:
        d[...] = m
        return
    rows, hp = _hptrs(h)
    m2 = np.ascontiguousarray(m, np.float32).reshape(-1, m.shape[-1])
    d2 = d.reshape(-1, d.shape[-1])
    _lib.load("spacetime").interpforward_nd_float(ctypes.addressof(hp), d2.ctypes.data, m2.ctypes.data, len(rows),
                                                  d2.shape[1], m2.shape[1], d2.shape[0], 1)


def interpadjoint(h, m, d):
    """coarse d (nslow, nd) -> fine m (nslow, nm), src/spacetime.jl:522-534."""
    if m.shape[-1] == d.shape[-1]:
        m[...] = d
        return
    rows, hp = _hptrs(h)
    d2 = np.ascontiguousarray(d, np.float32).reshape(-1, d.shape[-1])
    m2 = m.reshape(-1, m.shape[-1])
    _lib.load("spacetime").interpadjoint_nd_float(ctypes.addressof(hp), m2.ctypes.data, d2.ctypes.data, len(rows),
                                                  m2.shape[1], d2.shape[1], d2.shape[0], 1)


# ---- wavelets (src/wavelet.jl:150-181) ---------------------------------------------------

def ricker(f, t, a=1.0):
    """a (1 - 2 (pi f t)^2) exp(-(pi f t)^2), evaluated in float64 then converted."""
    t = np.asarray(t, np.float64)
    u = (np.pi * f * t) ** 2
    return (a * (1.0 - 2.0 * u) * np.exp(-u)).astype(np.float32)


def causal_ricker(f, t, a=1.0, tol=1e-4):
    """Ricker centred on mean(t), leading samples with |w| <= tol*a trimmed (shifted to t0)."""
    t = np.asarray(t, np.float64)
    w = ricker(f, t - t.mean(), a)
    idx = np.nonzero(np.abs(w) > tol * a)[0][0]
    out = np.zeros_like(w)
    out[:len(w) - idx] = w[idx:]
    return out
