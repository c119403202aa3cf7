"""Snapshot store, mirroring src/compressedio.jl (Compressor, compressedwrite,
compressedread!, comptype).

Raw modes -- C equal to the field type, or Float64 fields stored as Float32 -- write and read
the values exactly as the reference does (offset (ifield-1)*length*sizeof(P)).  The lossy
UInt32 mode, which the reference delegates to the un-vendored CvxCompress, is served by
libwavefd_codec (per-block max-abs scaling, 8-bit levels for scale >= 1/127, else 16-bit),
with the reference's per-field offsets/lengths bookkeeping for random access.
"""
import ctypes

import numpy as np

from . import _lib


def comptype(C, T):
    """src/compressedio.jl:138-148: returns (C, P)."""
    C = T if C is None else C
    if T == np.float64:
        assert C in (np.float64, np.float32, np.uint32)
    else:
        assert C in (np.float32, np.uint32)
    P = np.float64 if (T == np.float64 and C == np.float64) else np.float32
    return C, P


class Compressor:
    """src/compressedio.jl:1-16: per-field offsets/lengths (bytes) and the codec settings."""

    def __init__(self, T, P, C, n, b, scale, nfields, issubarray):
        self.T, self.P, self.C = T, P, C
        self.n = tuple(int(v) for v in n)
        self.b = tuple(int(v) for v in b)
        self.scale = float(scale)
        self.offsets = np.zeros(nfields, np.int64)
        self.lengths = np.zeros(nfields, np.int64)
        self.issubarray = issubarray
        self.bits = 8 if self.scale >= 1.0 / 127 else 16

    def open(self):
        return None

    def close(self):
        return None

    def _codec_geometry(self, shape):
        # Julia (n1 fastest): a Fortran-ordered array's shape is (n1, n2, n3)
        n = (ctypes.c_long * 3)(*(list(shape) + [1] * (3 - len(shape))))
        b = (ctypes.c_long * 3)(*(list(self.b) + [1] * (3 - len(self.b))))
        return n, b


class CompressedWriteError(IOError):
    """short write (src/compressedio.jl:33-37)"""


def _as_julia(field, rng):
    a = field if rng is None else field[tuple(slice(r.start, r.stop) for r in rng)]
    return np.asfortranarray(a)


def compressedwrite(io, comp, ifield, field, rng=None):
    """Write field (or field[rng]) as field number ifield (1-based)."""
    a = _as_julia(field, rng)
    if comp.C == np.uint32:
        lib = _lib.load("codec")
        src = np.asfortranarray(a, np.float32)
        n, b = comp._codec_geometry(src.shape)
        nw = lib.wavefd_codec_words(n, b, comp.bits)
        buf = np.empty(nw, np.uint32)
        got = lib.wavefd_codec_compress(src.ctypes.data, n, b, comp.bits, buf.ctypes.data)
        if got < 0:
            raise RuntimeError(_lib.last_error("codec"))
        data = buf[:got].tobytes()
        comp.lengths[ifield - 1] = len(data)
        nbytes = io.write(data)
        if nbytes != len(data):
            raise CompressedWriteError("wrote %d of %d" % (nbytes, len(data)))
        if ifield < len(comp.offsets):
            comp.offsets[ifield] = comp.offsets[ifield - 1] + comp.lengths[ifield - 1]
        return
    data = np.asfortranarray(a, comp.P).tobytes(order="F")
    nbytes = io.write(data)
    if nbytes != len(data):
        raise CompressedWriteError("wrote %d of %d" % (nbytes, len(data)))


def compressedread(io, comp, ifield, field, rng=None):
    """Read field number ifield into field (or field[rng]), src/compressedio.jl:92-136."""
    target = field if rng is None else field[tuple(slice(r.start, r.stop) for r in rng)]
    if comp.C == np.uint32:
        io.seek(int(comp.offsets[ifield - 1]))
        raw = io.read(int(comp.lengths[ifield - 1]))
        words = np.frombuffer(raw, np.uint32).copy()
        out = np.empty(target.shape, np.float32, order="F")
        if _lib.load("codec").wavefd_codec_decompress(words.ctypes.data, len(words), out.ctypes.data) != 0:
            raise RuntimeError(_lib.last_error("codec"))
        target[...] = out
        return
    nbytes = target.size * np.dtype(comp.P).itemsize
    io.seek((ifield - 1) * nbytes)
    vals = np.frombuffer(io.read(nbytes), comp.P).reshape(target.shape, order="F")
    target[...] = vals
