"""Learned snapshot codec (libwavefd_lcodec): a fixed 11-level block quantiser (3.59 bits per
value, 8.9x smaller than float32) and a learned convolutional decoder that removes part of the
quantisation error, trained on the GPU on snapshots the propagator generates.

The reference keeps Born/RTM snapshots exactly or through CvxCompress
(src/compressedio.jl:1-148); exact storage stays the default of ``born.born_shot`` and this
codec is the opt-in replacement (``snapshot_codec=``).  All arrays are device addresses; a
field is (nz, ny, nx) in Julia order, i.e. nz fastest.
"""
import ctypes

import numpy as np

from . import _lib, ginsu

F32 = np.float32


def _err():
    return _lib.last_error("lcodec")


LEVELS = (11, 5, 3)  # quantiser levels: 3.56, 2.5 or 1.63 bits per value


def code_words(shape, levels=11):
    """32-bit words of the code of a (nz, ny, nx) field"""
    nz, ny, nx = shape
    n = int(_lib.load("lcodec").wavefd_lcodec_code_words_l(nz, ny, nx, int(levels)))
    if n < 0:
        raise ValueError("quantiser levels: one of %s" % (LEVELS,))
    return n


def ratio(shape, levels=11):
    """float32 bytes / code bytes"""
    nz, ny, nx = shape
    return nz * ny * nx / code_words(shape, levels)


def encode(field, shape, code, stream=None, levels=11):
    """quantise the device field (nz, ny, nx) into the device buffer code (code_words words):
    8^3 blocks, a bfloat16 scale per block (its largest |x| rounded up), q = rint((levels - 1) x /
    (2 s))"""
    nz, ny, nx = shape
    if _lib.load("lcodec").wavefd_lcodec_encode_l(int(field), nz, ny, nx, int(levels), int(code), stream) != 0:
        raise RuntimeError(_err())


def dequantize(code, out, stream=None, shape=None, levels=11):
    """the plain decode (no network): out = 2 s d / (levels - 1); with the field's (nz, ny, nx)
    shape (and the code's levels) given, without reading the code's header (no host wait)"""
    lib = _lib.load("lcodec")
    if shape is None:
        rc = lib.wavefd_lcodec_dequantize(int(code), int(out), stream)
    else:
        rc = lib.wavefd_lcodec_dequantize_nl(int(code), *[int(v) for v in shape], int(levels), int(out), stream)
    if rc != 0:
        raise RuntimeError(_err())


def combine(n, ca, a, out, cb=0.0, b=None, cc=0.0, c=None, stream=None):
    """out = ca a + cb b + cc c on n floats (device addresses; b, c optional), in that order"""
    if _lib.load("lcodec").wavefd_lcodec_combine(int(n), float(ca), int(a), float(cb), None if b is None else int(b),
                                                 float(cc), None if c is None else int(c), int(out), stream) != 0:
        raise RuntimeError(_err())


def _opt(p):
    return None if p is None else int(p)


def predict_residual(n, coef, prev, x, pred, res, stream=None):
    """pred = sum coef[i] prev[i] (up to 3 terms, in order) and res = x - pred, in one pass"""
    c = list(coef) + [0.0] * (3 - len(coef))
    p = list(prev) + [None] * (3 - len(prev))
    if _lib.load("lcodec").wavefd_lcodec_predict_residual(int(n), float(c[0]), int(p[0]), float(c[1]), _opt(p[1]),
                                                          float(c[2]), _opt(p[2]), int(x), int(pred), int(res),
                                                          stream) != 0:
        raise RuntimeError(_err())


class LearnedCodec:
    """The quantiser with its learned decoder.  Parameters: per layer W[t][co][c] (27 x 16 x 16,
    t = (dx+1) 9 + (dy+1) 3 + (dz+1)) and 16 biases, three layers (conv1 reads channel 0 only,
    conv3 writes channel 0 only)."""

    def __init__(self, seed=0, params=None):
        self.lib = _lib.load("lcodec")
        self.net = self.lib.wavefd_lcodec_net_create()
        if not self.net:
            raise RuntimeError(_err())
        self.nparams = int(self.lib.wavefd_lcodec_nparams())
        if params is not None:
            self.params = params
        elif self.lib.wavefd_lcodec_net_init(self.net, seed) != 0:
            raise RuntimeError(_err())

    def close(self):
        if getattr(self, "net", None):
            self.lib.wavefd_lcodec_net_free(self.net)
            self.net = None

    def __del__(self):
        self.close()

    @property
    def params(self):
        p = np.zeros(self.nparams, F32)
        if self.lib.wavefd_lcodec_net_get(self.net, p.ctypes.data) != 0:
            raise RuntimeError(_err())
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, F32)
        if p.size != self.nparams:
            raise ValueError("expected %d parameters" % self.nparams)
        if self.lib.wavefd_lcodec_net_set(self.net, p.ctypes.data) != 0:
            raise RuntimeError(_err())

    def save(self, path):
        np.savez(path, params=self.params)

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            return cls(params=z["params"])

    # ---- network on volume batches (device arrays of npatch x X x Y x Z float32) ----------
    def forward(self, q, vol, out, norm=1.0, hs=None, stream=None):
        """out = q + norm f(q / norm) on a batch vol = (npatch, X, Y, Z) of volumes, z fastest; the
        correction clamped to +-hs (device array of half steps) when given"""
        n, X, Y, Z = vol
        if self.lib.wavefd_lcodec_forward(self.net, int(q), n, X, Y, Z, F32(norm), None if hs is None else int(hs),
                                          int(out), stream) != 0:
            raise RuntimeError(_err())

    def grad(self, q, target, vol, out, want_grad=False, norm=1.0, hs=None, stream=None):
        """loss (mean squared error of out = q + norm f(q / norm) against target) and, if asked,
        the gradient in the parameters (host array); the gradient also stays for ``adam``"""
        n, X, Y, Z = vol
        loss = ctypes.c_float(0.0)
        g = np.zeros(self.nparams, F32) if want_grad else None
        rc = self.lib.wavefd_lcodec_grad(self.net, int(q), int(target), n, X, Y, Z, F32(norm),
                                         None if hs is None else int(hs), int(out), ctypes.byref(loss),
                                         g.ctypes.data if want_grad else None, stream)
        if rc != 0:
            raise RuntimeError(_err())
        return float(loss.value), g

    def adam(self, lr, stream=None):
        if self.lib.wavefd_lcodec_adam(self.net, F32(lr), stream) != 0:
            raise RuntimeError(_err())

    # ---- codec ----------------------------------------------------------------------------
    def encode(self, field, shape, code, scratch1, scratch2, stream=None, levels=11):
        """quantise and record, per 8^3 block, whether the learned decode beats the plain one
        (scratch1, scratch2: device arrays of the field's size)"""
        nz, ny, nx = shape
        if self.lib.wavefd_lcodec_encode_learned_l(self.net, int(field), nz, ny, nx, int(levels), int(code),
                                                   int(scratch1), int(scratch2), stream) != 0:
            raise RuntimeError(_err())

    def decode(self, code, out, scratch, stream=None):
        """learned decode of a code into out (plain in the blocks the encoder chose so);
        scratch: a device array of the field's size"""
        if self.lib.wavefd_lcodec_decode(self.net, int(code), int(scratch), int(out), stream) != 0:
            raise RuntimeError(_err())

    def train(self, snapshots, shape, iters=1000, patch=32, batch=4, lr=1e-3, seed=0, log_every=0, levels=11):
        """Train the decoder on device snapshots (addresses of (nz, ny, nx) float32 fields): each
        step takes `batch` patch^3 cubes of random snapshots (half of them drawn where the energy
        is, half uniformly), their quantised-dequantised
        versions as input and the exact cubes as target, both divided by the snapshot's code norm
        (the decode's normalisation), the correction clamped to half a quantisation step as in
        the decode, and makes one Adam step.  Returns the per-step losses."""
        nz, ny, nx = shape
        p = min(patch, nz, ny, nx)
        n = nz * ny * nx
        cw = code_words(shape, levels)
        code = ginsu.DeviceBuffer(4 * cw)
        deq, half, inv = [], [], []
        for s in snapshots:  # the dequantised twin of every snapshot, its half steps, 1 / its norm
            d, h = ginsu.DeviceBuffer(4 * n), ginsu.DeviceBuffer(4 * n)
            encode(s, shape, code.ptr, levels=levels)
            dequantize(code.ptr, d.ptr)
            if self.lib.wavefd_lcodec_halfstep(code.ptr, h.ptr, None) != 0:
                raise RuntimeError(_err())
            g = float(self.lib.wavefd_lcodec_norm(code.ptr))
            if g <= 0:
                raise RuntimeError(_err())
            deq.append(d)
            half.append(h)
            inv.append(1.0 / g)
        code.free()
        # where the energy is: the relative error of a snapshot (and the image) is energy
        # weighted, so half of the patches are centred on an 8^3 block drawn with probability
        # proportional to its energy, the other half uniformly
        cdfs = []
        for sp in snapshots:
            x = np.zeros(n, F32)
            ginsu.copy(x.ctypes.data, sp, 4 * n)
            ginsu.sync()
            x = x.reshape(nx, ny, nz).astype(np.float64) ** 2
            bx, by, bz = -(-nx // 8), -(-ny // 8), -(-nz // 8)
            e = np.zeros((bx * 8, by * 8, bz * 8))
            e[:nx, :ny, :nz] = x
            e = e.reshape(bx, 8, by, 8, bz, 8).sum(axis=(1, 3, 5)).ravel()
            cdfs.append((np.cumsum(e) / max(e.sum(), 1e-300), (bx, by, bz)))
        per = p * p * p
        qb, tb, hb, ob = (ginsu.DeviceBuffer(4 * per * batch) for _ in range(4))
        rng = np.random.default_rng(seed)
        lib = self.lib
        losses = []

        def origin(k):
            if rng.random() < 0.5:
                return np.array([rng.integers(0, nz - p + 1), rng.integers(0, ny - p + 1),
                                 rng.integers(0, nx - p + 1)], np.int64)
            cdf, (bx, by, bz) = cdfs[k]
            blk = min(int(np.searchsorted(cdf, rng.random())), len(cdf) - 1)
            cx, cy, cz = blk // (by * bz) * 8 + 4, (blk // bz) % by * 8 + 4, blk % bz * 8 + 4
            o = [int(np.clip(c - p // 2 + rng.integers(-p // 4, p // 4 + 1), 0, m - p))
                 for c, m in ((cz, nz), (cy, ny), (cx, nx))]
            return np.array(o, np.int64)

        try:
            for it in range(iters):
                for b in range(batch):
                    k = int(rng.integers(len(snapshots)))
                    org = origin(k)
                    for src, dst in ((deq[k].ptr, qb.ptr), (snapshots[k], tb.ptr), (half[k].ptr, hb.ptr)):
                        if lib.wavefd_lcodec_gather(int(src), nz, ny, nx, org.ctypes.data, 1, p, F32(inv[k]),
                                                    dst + 4 * per * b, None) != 0:
                            raise RuntimeError(_err())
                loss, _ = self.grad(qb.ptr, tb.ptr, (batch, p, p, p), ob.ptr, hs=hb.ptr)
                self.adam(lr)
                losses.append(loss)
                if log_every and it % log_every == 0:
                    print("lcodec train step %d loss %.4e" % (it, loss), flush=True)
        finally:
            for d in deq + half + [qb, tb, hb, ob]:
                d.free()
        return np.array(losses)
