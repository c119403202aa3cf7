"""Ginsu windows: per-shot sub-grids of a large model, and the shared image they add into.

The reference leaves windowing to its callers (src/WaveFD.jl:52); these rules are this
package's own, integer-exact and documented here:

* a big grid is Julia-ordered (nz, ny, nx) with spacings (dz, dy, dx) and origin (z0, y0, x0);
  cell (iz, iy, ix) sits at (z0 + iz dz, y0 + iy dy, x0 + ix dx);
* ``Window(big_shape, origin, size)`` is the box [oz, oz+snz) x [oy, oy+sny) x [ox, ox+snx),
  wholly inside the big grid;
* ``window_for_shot`` covers the source and receivers plus a horizontal aperture and a pad of
  cells (for the absorbing sponge): along x, lo = floor((min x - aperture - x0) / dx) - pad and
  hi = ceil((max x + aperture - x0) / dx) + 1 + pad, clipped to [0, nx); the same along y;
  along z the window runs from the top of the grid (the free surface) to the deepest of the
  grid bottom and ``zmax``; positions are float64 and the floors/ceils are exact integer
  arithmetic on them;
* ``Window.interior(nsponge, freesurface)`` is the box of window cells that the shared image
  receives: the window minus its sponge, except on faces that are also faces of the big grid
  (there the sponge is the big model's own) and except the free-surface top.

Extraction and accumulation run on the device through libwavefd_ginsu (include/wavefd_ginsu.h).
"""
import math

import numpy as np

from . import _lib


class Window:
    def __init__(self, big_shape, origin, size):
        self.big = tuple(int(n) for n in big_shape)
        self.origin = tuple(int(o) for o in origin)
        self.size = tuple(int(s) for s in size)
        for n, o, s in zip(self.big, self.origin, self.size):
            if s <= 0 or o < 0 or o + s > n:
                raise ValueError("window %s + %s outside the big grid %s" % (self.origin, self.size, self.big))

    @property
    def shape(self):
        """(nz, ny, nx) of the window"""
        return self.size

    @property
    def ncells(self):
        return self.size[0] * self.size[1] * self.size[2]

    def interior(self, nsponge, freesurface=True):
        """((lz0, lz1), (ly0, ly1), (lx0, lx1)) window-relative box that the shared image receives."""
        box = []
        for axis, (n, o, s) in enumerate(zip(self.big, self.origin, self.size)):
            lo = 0 if o == 0 else nsponge
            hi = s if o + s == n else s - nsponge
            if axis == 0 and freesurface and o == 0:
                lo = 0
            box.append((min(lo, s), max(min(hi, s), min(lo, s))))
        return tuple(box)

    def __repr__(self):
        return "Window(origin=%s, size=%s, big=%s)" % (self.origin, self.size, self.big)


def _lo(v, v0, d, pad):
    return int(math.floor((v - v0) / d)) - pad


def _hi(v, v0, d, pad):
    return int(math.ceil((v - v0) / d)) + 1 + pad


def window_for_shot(big_shape, spacing, origin, src, rcv_z, rcv_y, rcv_x, aperture=(0.0, 0.0), pad=0, zmax=None):
    """Window covering the shot: src = (z, y, x); receivers as arrays; aperture = (ay, ax) in
    length units; pad in cells on every clipped face; zmax: deepest z the window must reach
    (default: the whole grid depth)."""
    nz, ny, nx = big_shape
    dz, dy, dx = spacing
    z0, y0, x0 = origin
    ys = np.concatenate([[src[1]], np.asarray(rcv_y, np.float64)])
    xs = np.concatenate([[src[2]], np.asarray(rcv_x, np.float64)])
    ay, ax = aperture
    ylo = max(0, _lo(ys.min() - ay, y0, dy, pad))
    yhi = min(ny, _hi(ys.max() + ay, y0, dy, pad))
    xlo = max(0, _lo(xs.min() - ax, x0, dx, pad))
    xhi = min(nx, _hi(xs.max() + ax, x0, dx, pad))
    zhi = nz if zmax is None else min(nz, _hi(zmax, z0, dz, pad))
    return Window(big_shape, (0, ylo, xlo), (zhi, yhi - ylo, xhi - xlo))


def extract(big, window, sub, stream=None):
    """sub (device, window-sized) = big (device)[window], bit-exact, ordered on stream."""
    nz, ny, nx = window.big
    oz, oy, ox = window.origin
    snz, sny, snx = window.size
    lib = _lib.load("ginsu")
    if lib.wavefd_ginsu_extract(int(big), nz, ny, nx, oz, oy, ox, int(sub), snz, sny, snx, stream) != 0:
        raise RuntimeError(_lib.last_error("ginsu"))


def accumulate(big, window, sub, box=None, stream=None):
    """big (device)[window][box] += sub (device)[box]; box defaults to the whole window."""
    nz, ny, nx = window.big
    oz, oy, ox = window.origin
    snz, sny, snx = window.size
    if box is None:
        box = ((0, snz), (0, sny), (0, snx))
    (lz0, lz1), (ly0, ly1), (lx0, lx1) = box
    lib = _lib.load("ginsu")
    if lib.wavefd_ginsu_accumulate(int(big), nz, ny, nx, oz, oy, ox, int(sub), snz, sny, snx, lz0, lz1, ly0, ly1,
                                   lx0, lx1, stream) != 0:
        raise RuntimeError(_lib.last_error("ginsu"))


class DeviceBuffer:
    """A device allocation (through libwavefd_ginsu) for drivers that hold no allocator."""

    def __init__(self, nbytes):
        self.lib = _lib.load("ginsu")
        self.nbytes = int(nbytes)
        self.ptr = self.lib.wavefd_device_alloc(self.nbytes)
        if not self.ptr:
            raise MemoryError(_lib.last_error("ginsu"))

    def zero(self, stream=None, offset=0, nbytes=None):
        n = self.nbytes - offset if nbytes is None else nbytes
        if self.lib.wavefd_device_zero(self.ptr + offset, n, stream) != 0:
            raise RuntimeError(_lib.last_error("ginsu"))

    def upload(self, host, offset=0, stream=None):
        host = np.ascontiguousarray(host)
        if self.lib.wavefd_device_copy(self.ptr + offset, host.ctypes.data, host.nbytes, stream) != 0:
            raise RuntimeError(_lib.last_error("ginsu"))
        self.sync(stream)

    def download(self, host, offset=0, stream=None):
        if not host.flags["C_CONTIGUOUS"] and not host.flags["F_CONTIGUOUS"]:
            raise ValueError("host array must be contiguous")
        if self.lib.wavefd_device_copy(host.ctypes.data, self.ptr + offset, host.nbytes, stream) != 0:
            raise RuntimeError(_lib.last_error("ginsu"))
        self.sync(stream)
        return host

    def sync(self, stream=None):
        if self.lib.wavefd_device_sync(stream) != 0:
            raise RuntimeError(_lib.last_error("ginsu"))

    def free(self):
        if self.ptr:
            self.lib.wavefd_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy(dst, src, nbytes, stream=None):
    """stream-ordered copy between device (or host) addresses"""
    lib = _lib.load("ginsu")
    if lib.wavefd_device_copy(int(dst), int(src), int(nbytes), stream) != 0:
        raise RuntimeError(_lib.last_error("ginsu"))


def sync(stream=None):
    """wait for the work queued on a stream (None: the null stream)"""
    lib = _lib.load("ginsu")
    if lib.wavefd_device_sync(stream) != 0:
        raise RuntimeError(_lib.last_error("ginsu"))


def zero(dst, nbytes, stream=None):
    lib = _lib.load("ginsu")
    if lib.wavefd_device_zero(int(dst), int(nbytes), stream) != 0:
        raise RuntimeError(_lib.last_error("ginsu"))
