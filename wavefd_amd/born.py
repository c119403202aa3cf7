"""Born modelling and imaging of one shot on a ginsu window, on the device.

The reference ships the single-step pieces (the propagators, ``forwardBornInjection_V`` and
``adjointBornAccumulation_V``, src/prop3DAcoIsoDenQ_DEO2_FDTD.h:305-334 and :518-546, and their
VTI twins) and leaves the time loops to its callers (JetPackWaveFD, out of tree).  This driver
writes those loops against this package's C-ABI, all state in HBM:

1. background forward: per step ``TimeStep``, inject the source, ``ScaleSpatialDerivatives``;
   DP_n = v^2/b L(p_{n-1}) (and DM_n) saved to a device snapshot buffer (by default through the
   fused ``TimeStepSaveDP``, and likewise the fused pairs of passes 2 and 3);
2. Born forward (the linearisation J in dv): per step ``TimeStep``, ``ForwardBornInjection_V``
   with DP_n (p += 2 dt^2 dv/v DP_n into the new field), extract the receivers: d = J dv;
3. adjoint (J^T): the forward step is self-adjoint in the b/(dt^2 v^2)-weighted inner product,
   so J^T d runs the same step on the time-reversed data injected at the receivers with
   weights dt^2 v^2 / b, and images with ``AdjointBornAccumulation_V`` (g += 2b/v^3 DP_n lam_n,
   which carries the weight b/(dt^2 v^2) of the sources' side).  That call images the field
   in pOld, i.e. the previous step's field, so DP_{nt+2-m} is used after reversed step m,
   and one extra step images DP_1;
4. the window's image is added into the shared image over the window's interior
   (``Window.interior``).

For Iso, passes 1 and 2 run as one pass over two wavefields (``TimeStepBornCoupled_V``): the
model is read once per step for both, and the Born source uses DP_n as it is formed.

``tests/test_gpu_born_pipeline.py`` checks <J dv, d> = <dv, J^T d> and J against a central
difference of the full modelling.
"""
import ctypes
import time

import numpy as np

from . import ginsu, prop3d, spacetime

F32 = np.float32
_CLS = {"iso": prop3d.Prop3DAcoIsoDenQ_DEO2_FDTD, "vti": prop3d.Prop3DAcoVTIDenQ_DEO2_FDTD}
MODEL_FIELDS = {"iso": ("v", "b"), "vti": ("v", "b", "eps", "eta", "f")}


def _shift(points_z, points_y, points_x, window, spacing):
    oz, oy, ox = window.origin
    dz, dy, dx = spacing
    return (np.asarray(points_z, np.float64) - oz * dz, np.asarray(points_y, np.float64) - oy * dy,
            np.asarray(points_x, np.float64) - ox * dx)


def _gather(field_dev, pts, nsteps=1):
    """values of a device field at the taps of pts (one per tap, in tap order), via the
    device extraction path of libspacetime"""
    taps = pts.copy()
    taps["ir"] = np.arange(1, len(taps) + 1)
    taps["c"] = 1.0
    buf = ginsu.DeviceBuffer(max(4, 4 * len(taps)))
    buf.zero()
    buf.sync()
    spacetime.extractdata(_DevArray(buf.ptr, (len(taps), 1)), field_dev, 1, [taps])
    out = np.zeros(len(taps), F32)
    buf.download(out)
    buf.free()
    return out


class _DevArray:
    """a device address with the (nr, nt) shape that libspacetime's wrappers read nt from"""

    def __init__(self, ptr, shape):
        self.ptr, self.shape = ptr, shape

    def data_ptr(self):
        return self.ptr


class _Slice:
    """a piece of a device buffer: the address and size an exact store needs (not owned)"""

    def __init__(self, ptr, nbytes):
        self.ptr, self.nbytes = ptr, nbytes


class _ExactStore:
    """DP_n (DM_n) kept exactly: one float32 field per step and wavefield"""

    def __init__(self, ncell, nt, nf, buffer=None):
        self.ncell, self.nf = ncell, nf
        self.own = buffer is None
        self.buf = ginsu.DeviceBuffer(4 * ncell * nt * nf) if buffer is None else buffer
        if self.buf.nbytes < 4 * ncell * nt * nf:
            raise ValueError("snapshot buffer too small")

    def slot(self, kind, n):
        return self.buf.ptr + 4 * self.ncell * ((n - 1) * self.nf + kind)

    def begin(self, n):  # forward step n is about to run
        pass

    def out(self, kind, n):  # where the step writes DP_n
        return self.slot(kind, n)

    def commit(self, kind, n, stream):
        pass

    def get(self, kind, n, stream):  # where DP_n can be read
        return self.slot(kind, n)

    def free(self):
        if self.own:
            self.buf.free()


class _CodecStore:
    """DP_n (DM_n) kept as learned-codec codes (lcodec): written to a scratch field, encoded;
    read back through the learned decoder (or the plain dequantisation, learned=False)"""

    def __init__(self, shape, nt, nf, codec, learned=True):
        from . import lcodec
        self.lc, self.codec, self.learned, self.shape, self.nf = lcodec, codec, learned, shape, nf
        self.words = lcodec.code_words(shape)
        ncell = int(np.prod(shape))
        self.codes = ginsu.DeviceBuffer(4 * self.words * nt * nf)
        self.scratch = [ginsu.DeviceBuffer(4 * ncell) for _ in range(nf)]
        self.tmp = ginsu.DeviceBuffer(4 * ncell)
        self.tmp2 = ginsu.DeviceBuffer(4 * ncell) if learned else None

    def code(self, kind, n):
        return self.codes.ptr + 4 * self.words * ((n - 1) * self.nf + kind)

    def begin(self, n):
        pass

    def out(self, kind, n):
        return self.scratch[kind].ptr

    def commit(self, kind, n, stream):
        if self.learned:  # with the encoder's per-block choice of the learned decode
            self.codec.encode(self.scratch[kind].ptr, self.shape, self.code(kind, n), self.tmp.ptr, self.tmp2.ptr,
                              stream)
        else:
            self.lc.encode(self.scratch[kind].ptr, self.shape, self.code(kind, n), stream)

    def get(self, kind, n, stream):
        if self.learned:
            self.codec.decode(self.code(kind, n), self.scratch[kind].ptr, self.tmp.ptr, stream)
        else:
            self.lc.dequantize(self.code(kind, n), self.scratch[kind].ptr, stream)
        return self.scratch[kind].ptr

    def free(self):
        for b in [self.codes, self.tmp] + self.scratch + ([self.tmp2] if self.tmp2 else []):
            b.free()


def checkpoint_fields(nt, nf, k):
    """Fields of one window a checkpointed store holds for nt steps, checkpoints every k steps:
    2 nf per checkpoint (ceil(nt / k) of them), k nf for the segment at hand, and the second
    propagator's state fields (Iso: v b Q pCur pOld pSpace; VTI: 12).  Exact storage: nt nf."""
    nseg = (nt + k - 1) // k
    return 2 * nf * nseg + min(k, nt) * nf + (6 if nf == 1 else 12)


class _CheckpointStore:
    """DP_n (DM_n) recomputed instead of stored.  The background wavefields (pCur, pOld[, mCur,
    mOld]) are copied every k steps of the forward pass (checkpoints: the state before steps 1,
    k+1, 2k+1, ...).  When a pass asks for DP_n of a segment that is not at hand, the segment's
    steps run again from its checkpoint on a second propagator holding the same model, Q and
    source, with the forward pass's own step and injection calls; the kernels are
    deterministic, so every recomputed DP_n is bit-identical to the one the forward pass formed
    and the image equals that of exact storage bit for bit.  The forward pass writes the last
    segment's DP_n straight into the segment buffer, so the adjoint's first segment needs no
    recomputation.  Costs: one more background pass for the adjoint (and one for a separate Born
    forward pass, i.e. VTI); memory: ``checkpoint_fields`` instead of nt nf fields.  The second
    propagator runs on its own stream; the main stream is drained before a segment is
    overwritten and the second stream before its DP_n are read."""

    def __init__(self, run, k, fused):
        if int(k) < 1:
            raise ValueError("checkpoint interval must be >= 1")
        self.run, self.k, self.fused = run, int(k), fused
        self.nf, self.nt, self.ncell = run.nf, run.nt, run.ncell
        self.nseg = (self.nt + self.k - 1) // self.k
        self.seglen = min(self.k, self.nt)
        self.ckpt = ginsu.DeviceBuffer(max(4, 4 * self.ncell * 2 * self.nf * self.nseg))
        self.seg = ginsu.DeviceBuffer(max(4, 4 * self.ncell * self.seglen * self.nf))
        self.scratch = ginsu.DeviceBuffer(4 * self.ncell * self.nf)
        self.cached = None  # the segment whose DP_n the segment buffer holds
        self.prop2 = None

    def _names(self):
        return ("pCur", "pOld") + (("mCur", "mOld") if self.nf == 2 else ())

    def _ckpt(self, j, i):
        return self.ckpt.ptr + 4 * self.ncell * (j * 2 * self.nf + i)

    def _slot(self, kind, n):
        j = (n - 1) // self.k
        return self.seg.ptr + 4 * self.ncell * ((n - 1 - j * self.k) * self.nf + kind)

    def begin(self, n):
        if (n - 1) % self.k == 0:
            r = self.run
            for i, name in enumerate(self._names()):
                ginsu.copy(self._ckpt((n - 1) // self.k, i), prop3d.device_ptr(r.prop, name), 4 * self.ncell,
                           r.stream)

    def out(self, kind, n):
        j = (n - 1) // self.k
        if j == self.nseg - 1:
            self.cached = j
            return self._slot(kind, n)
        return self.scratch.ptr + 4 * self.ncell * kind

    def commit(self, kind, n, stream):
        pass

    def _second(self):
        r = self.run
        p2 = r.new_prop()
        # the model; Q needs no copy: the same construction parameters give the same profile
        # (and handing out Q's address would end the fused step's table lookup of it)
        for name in MODEL_FIELDS[r.physics]:
            ginsu.copy(prop3d.device_ptr(p2, name), prop3d.device_ptr(r.prop, name), 4 * self.ncell,
                       p2._lib.GetStream(p2.p))
        prop3d.set_sources(p2, [r.src_pts], r.shot.wavelet, fields=1 if self.nf == 1 else 3)
        prop3d.set_receivers(p2, [], self.nt)
        return p2

    def _recompute(self, j, first=False):
        """the steps of segment j on the second propagator, their DP_n into the segment buffer:
        from checkpoint j, or (first) from the second propagator's own state, which is then
        saved as checkpoint j -- the background pass itself, segment by segment"""
        r = self.run
        prop3d.synchronize(r.prop)  # the checkpoints are written and the old segment is read
        if self.prop2 is None:
            self.prop2 = self._second()
        p2 = self.prop2
        lib, h, stream = p2._lib, p2.p, p2._lib.GetStream(p2.p)
        for i, name in enumerate(self._names()):
            if first:
                ginsu.copy(self._ckpt(j, i), prop3d.device_ptr(p2, name), 4 * self.ncell, stream)
            else:
                ginsu.copy(prop3d.device_ptr(p2, name), self._ckpt(j, i), 4 * self.ncell, stream)
        for it in range(j * self.k + 1, min((j + 1) * self.k, self.nt) + 1):
            slots = [self._slot(kind, it) for kind in range(self.nf)]
            if self.fused:
                lib.TimeStepSaveDP(h, *slots)
                lib.InjectRegistered(h, it)
            else:
                lib.TimeStepInjectExtract(h, it)
                lib.ScaleSpatialDerivatives(h)
                ginsu.copy(slots[0], prop3d.device_ptr(p2, "pSpace"), 4 * self.ncell, stream)
                if self.nf == 2:
                    ginsu.copy(slots[1], prop3d.device_ptr(p2, "mSpace"), 4 * self.ncell, stream)
        p2._check()
        prop3d.synchronize(p2)
        self.cached = j

    def get(self, kind, n, stream):
        j = (n - 1) // self.k
        if self.cached != j:
            self._recompute(j)
        return self._slot(kind, n)

    def background_segment(self, j):
        """pass 1 for segment j on the second propagator (which starts from zero fields, like the
        main one), checkpointing it; a separate Born pass then reads the segment buffer"""
        self._recompute(j, first=True)

    def free(self):
        for b in (self.ckpt, self.seg, self.scratch):
            b.free()
        if self.prop2 is not None:
            prop3d.free(self.prop2)
            self.prop2 = None


class _PredictiveStore:
    """DP_n (DM_n) kept by closed-loop temporal prediction.  The frames fall into segments of k;
    the first frame of a segment (a keyframe) is kept by the block codec (libwavefd_codec, 16 or
    8 bits, error <= block max / 65534 or / 254), every other frame as the 11-level block quantiser's code
    (lcodec, 3.59 bits per value) of its residual against a polynomial extrapolation of the
    frames reconstructed before it in the segment: D_{n-1}; 2 D_{n-1} - D_{n-2}; 3 D_{n-1} -
    3 D_{n-2} + D_{n-3} (up to `order` of them).  The encoder predicts from its own
    reconstructions, so each frame's error is its own residual's quantisation error, at most a
    tenth of the residual's block maximum, and nothing accumulates.  A pass that asks for a frame
    of another segment decodes that segment into the segment buffer (keyframe, then prediction +
    decoded residual with the same kernel calls as the encoder, so the reconstructions are the
    encoder's bit for bit).

    Stored bytes: ceil(nt / k) keyframes of about 2 (16-bit) or 1 (8-bit) B per cell and
    nt - ceil(nt / k) residual codes of 0.45 B per cell (``stored_bytes``; k = 48: 8.2x fewer than
    float32); working fields: the segment buffer (k per wavefield) and order + 2 per wavefield
    for the encoder."""

    COEF = {1: (1.0,), 2: (2.0, -1.0), 3: (3.0, -3.0, 1.0)}

    def __init__(self, shape, nt, nf, k=48, order=3, keybits=16, codec=None):
        from . import lcodec, _lib
        if int(k) < 2 or int(order) not in self.COEF or int(keybits) not in (8, 16):
            raise ValueError("predictive store: segments of at least 2 frames, order 1, 2 or 3, 8- or 16-bit "
                             "keyframes")
        self.keybits = int(keybits)
        self.lc, self.shape, self.nt, self.nf, self.k, self.order = lcodec, shape, nt, nf, int(k), int(order)
        self.ncell = ncell = int(np.prod(shape))
        self.cod = _lib.load("codec")
        nz, ny, nx = shape
        self.n3 = (ctypes.c_long * 3)(nz, ny, nx)
        self.b3 = (ctypes.c_long * 3)(8, 8, 8)
        self.kwords = int(self.cod.wavefd_codec_words(self.n3, self.b3, self.keybits))
        self.words = lcodec.code_words(shape)
        self.nseg = (nt + self.k - 1) // self.k
        self.keys = ginsu.DeviceBuffer(4 * self.kwords * self.nseg * nf)
        self.codes = ginsu.DeviceBuffer(max(4, 4 * self.words * (nt - self.nseg) * nf))
        # encoder state per wavefield: the frame the step writes, the last `order`
        # reconstructions (newest first) and the prediction
        self.x = [ginsu.DeviceBuffer(4 * ncell) for _ in range(nf)]
        self.rec = [[ginsu.DeviceBuffer(4 * ncell) for _ in range(self.order)] for _ in range(nf)]
        self.pred = [ginsu.DeviceBuffer(4 * ncell) for _ in range(nf)]
        self.tmp = ginsu.DeviceBuffer(4 * ncell)
        self.seg = ginsu.DeviceBuffer(4 * ncell * min(self.k, nt) * nf)
        self.cached = None
        # optional learned residual decoder (lcodec.LearnedCodec, trained on residual frames):
        # its encode (with the per-block choice) and decode replace the plain quantiser's on both
        # sides, so the reconstructions still agree bit for bit
        self.codec = codec
        self.scr = [ginsu.DeviceBuffer(4 * ncell) for _ in range(2)] if codec is not None else []

    def _code_residual(self, code, stream):
        """quantise the residual in tmp into code, then leave its decode in tmp"""
        if self.codec is None:
            self.lc.encode(self.tmp.ptr, self.shape, code, stream)
            self.lc.dequantize(code, self.tmp.ptr, stream, self.shape)
        else:
            self.codec.encode(self.tmp.ptr, self.shape, code, self.scr[0].ptr, self.scr[1].ptr, stream)
            self.codec.decode(code, self.tmp.ptr, self.scr[0].ptr, stream)

    def _decode_residual(self, code, stream):
        if self.codec is None:
            self.lc.dequantize(code, self.tmp.ptr, stream, self.shape)
        else:
            self.codec.decode(code, self.tmp.ptr, self.scr[0].ptr, stream)

    def stored_bytes(self):
        return 4 * (self.kwords * self.nseg + self.words * (self.nt - self.nseg)) * self.nf

    def _key(self, kind, j):
        return self.keys.ptr + 4 * self.kwords * (j * self.nf + kind)

    def _code(self, kind, n):
        j, m = divmod(n - 1, self.k)
        return self.codes.ptr + 4 * self.words * ((j * (self.k - 1) + m - 1) * self.nf + kind)

    def _slot(self, kind, n):
        m = (n - 1) % self.k
        return self.seg.ptr + 4 * self.ncell * (m * self.nf + kind)

    def _predict(self, prev, out, stream):
        """the extrapolation of the reconstructions prev (newest first) into out; with one
        previous frame the prediction is that frame itself (returned, out untouched)"""
        c = self.COEF[len(prev)]
        if len(prev) == 1:
            return prev[0]
        self.lc.combine(self.ncell, c[0], prev[0], out, c[1], prev[1], c[2] if len(c) > 2 else 0.0,
                        prev[2] if len(c) > 2 else None, stream=stream)
        return out

    def _keyframe_decode(self, kind, j, dst, stream):
        ginsu.sync(stream)  # the block codec works on the null stream
        if self.cod.wavefd_codec_decompress(self._key(kind, j), self.kwords, dst) != 0:
            raise RuntimeError(_lib_error("codec"))

    def begin(self, n):
        pass

    def out(self, kind, n):  # where the step writes DP_n
        return self.x[kind].ptr

    def commit(self, kind, n, stream):
        lc, nc = self.lc, self.ncell
        m = (n - 1) % self.k
        x = self.x[kind].ptr
        rec = self.rec[kind]
        if m == 0:
            ginsu.sync(stream)
            if self.cod.wavefd_codec_compress(x, self.n3, self.b3, self.keybits, self._key(kind, (n - 1) // self.k)) < 0:
                raise RuntimeError(_lib_error("codec"))
            self._keyframe_decode(kind, (n - 1) // self.k, rec[-1].ptr, stream)
        elif self.codec is None:
            # prediction and residual in one pass; the reconstruction (prediction + decoded residual)
            # replaces the oldest one (no longer needed: the prediction is formed)
            q = min(m, self.order)
            pred = self.pred[kind].ptr
            lc.predict_residual(nc, self.COEF[q], [b.ptr for b in rec[:q]], x, pred, self.tmp.ptr, stream)
            lc.encode(self.tmp.ptr, self.shape, self._code(kind, n), stream)
            lc.dequantize(self._code(kind, n), self.tmp.ptr, stream, self.shape)
            lc.combine(nc, 1.0, pred, rec[-1].ptr, 1.0, self.tmp.ptr, stream=stream)
        else:
            q = min(m, self.order)
            pred = self._predict([b.ptr for b in rec[:q]], self.pred[kind].ptr, stream)
            lc.combine(nc, 1.0, x, self.tmp.ptr, -1.0, pred, stream=stream)  # residual
            self._code_residual(self._code(kind, n), stream)
            lc.combine(nc, 1.0, pred, rec[-1].ptr, 1.0, self.tmp.ptr, stream=stream)
        self.rec[kind] = [rec[-1]] + rec[:-1]

    def _decode_segment(self, j, stream):
        n0 = j * self.k + 1
        for kind in range(self.nf):
            self._keyframe_decode(kind, j, self._slot(kind, n0), stream)
            for n in range(n0 + 1, min(n0 + self.k, self.nt + 1)):
                q = min(n - n0, self.order)
                out = self._slot(kind, n)
                prev = [self._slot(kind, n - i) for i in range(1, q + 1)]
                pred = self._predict(prev, out, stream)
                self._decode_residual(self._code(kind, n), stream)
                self.lc.combine(self.ncell, 1.0, pred, out, 1.0, self.tmp.ptr, stream=stream)
        self.cached = j

    def get(self, kind, n, stream):
        j = (n - 1) // self.k
        if self.cached != j:
            self._decode_segment(j, stream)
        return self._slot(kind, n)

    def recon(self, kind):
        """the encoder's newest reconstruction (device address)"""
        return self.rec[kind][0].ptr

    def free(self):
        for b in ([self.keys, self.codes, self.tmp, self.seg] + self.x + self.pred + self.scr +
                  [r for rs in self.rec for r in rs]):
            b.free()


class _WaveStore:
    """DP_n (DM_n) kept by closed-loop prediction with the propagator itself.

    DP_n = (v^2 / b) L(p_n) is linear in p_n, and p_n = step(p_{n-1}, p_{n-2}) + w_n S (the
    registered source taps S scaled by the wavelet), so where Q is locally constant DP obeys the
    same leapfrog as the wavefield plus the source's image: DP_{n+1} = dt^2 (v^2/b) L(DP_n)
    - g (DP_n - DP_{n-1}) - DP_{n-1} + 2 DP_n + w_n G with G = (v^2/b) L(S) -- one TimeStep
    applied to (DP_n, DP_{n-1}) (for VTI to the pair (DP, DM), the coupled system being linear in
    (p, m) too) and one scaled add.  The store runs that step on a second propagator with the
    window's model and Q, whose pCur/pOld (mCur/mOld) buffers hold the last two reconstructions,
    forms G once per shot with the same propagator, and codes only the residual of each frame
    against the prediction (the 11-level block quantiser, 3.59 bits per value).  The prediction
    misses only the sponge's term (where Q varies), and the encoder predicts from its own
    reconstructions: an error that a frame carries is propagated into the next prediction and
    coded away with that frame's residual, so errors shrink from frame to frame instead of
    persisting.  Frames fall into segments of k; the
    first two frames of a segment are keyframes (the block codec, `keybits` bits).  Decoding a
    segment repeats the encoder's kernel calls (keyframe decodes, steps, dequantisation, the same
    combination), so its reconstructions are the encoder's bit for bit.

    Stored bytes: 2 keyframes and k - 2 residual codes per segment and wavefield (k = 96 with
    16-bit keyframes: 8.3x fewer than float32; measured on VTI 512^3, nt = 400: Born image within
    6.5e-6 of exact storage at k = 48, 1.3e-3 with 8-bit keyframes, whose own error then
    dominates); working fields: the segment buffer (k per wavefield), the frame the step writes,
    the residual, the source image G and the second propagator."""

    def __init__(self, run, k=96, keybits=16, levels=11, codec=None):
        from . import lcodec, _lib
        if int(k) < 3 or int(keybits) not in (8, 16) or int(levels) not in lcodec.LEVELS:
            raise ValueError("wave-predicted store: segments of at least 3 frames, 8- or 16-bit keyframes, "
                             "residuals quantised to %s levels" % (lcodec.LEVELS,))
        self.run, self.lc, self.k, self.keybits, self.levels = run, lcodec, int(k), int(keybits), int(levels)
        self.shape, self.nt, self.nf = run.window.size, run.nt, run.nf
        self.ncell = ncell = run.ncell
        self.cod = _lib.load("codec")
        nz, ny, nx = self.shape
        self.n3 = (ctypes.c_long * 3)(nz, ny, nx)
        self.b3 = (ctypes.c_long * 3)(8, 8, 8)
        self.kwords = int(self.cod.wavefd_codec_words(self.n3, self.b3, self.keybits))
        self.words = lcodec.code_words(self.shape, self.levels)
        self.nseg = (self.nt + self.k - 1) // self.k
        self.nkey = sum(min(2, self.nt - j * self.k) for j in range(self.nseg))  # keyframes per wavefield
        self.keys = ginsu.DeviceBuffer(4 * self.kwords * self.nkey * self.nf)
        self.codes = ginsu.DeviceBuffer(max(4, 4 * self.words * (self.nt - self.nkey) * self.nf))
        self.x = [ginsu.DeviceBuffer(4 * ncell) for _ in range(self.nf)]
        self.tmp = ginsu.DeviceBuffer(4 * ncell)
        self.seg = ginsu.DeviceBuffer(4 * ncell * min(self.k, self.nt) * self.nf)
        self.cached = None
        # optional learned residual decoder (lcodec.LearnedCodec trained on this store's residual
        # frames at these levels): its encode (with the per-block choice) and decode replace the
        # plain quantiser's on both sides, so the reconstructions still agree bit for bit
        self.codec = codec
        self.scr = [ginsu.DeviceBuffer(4 * ncell) for _ in range(2)] if codec is not None else []
        # the predictor: a propagator with the window's model; Q needs no copy (the same
        # construction parameters give the same profile)
        p2 = run.new_prop()
        for name in MODEL_FIELDS[run.physics]:
            ginsu.copy(prop3d.device_ptr(p2, name), prop3d.device_ptr(run.prop, name), 4 * ncell,
                       p2._lib.GetStream(p2.p))
        prop3d.synchronize(p2)
        self.p2 = p2
        self.cur = ("pCur", "mCur")
        self.old = ("pOld", "mOld")
        # G = (v^2/b) L(S): the unit taps in the current field(s), one step, the scaled spatial part
        self.wavelet = np.asarray(run.shot.wavelet, F32).reshape(-1)
        self.g = [ginsu.DeviceBuffer(4 * ncell) for _ in range(self.nf)]
        names = ("pCur", "pOld") + (("mCur", "mOld") if self.nf == 2 else ())
        for name in names:
            ginsu.zero(self._ptr(name), 4 * ncell, None)
        unit = np.ones((1, 1), F32)
        for kind in range(self.nf):
            spacetime.injectdata(self._ptr(self.cur[kind]), [run.src_pts], unit, 1)
        lib = p2._lib
        lib.TimeStep(p2.p)
        lib.ScaleSpatialDerivatives(p2.p)
        p2._check()
        prop3d.synchronize(p2)
        for kind, name in enumerate(("pSpace", "mSpace")[:self.nf]):
            ginsu.copy(self.g[kind].ptr, self._ptr(name), 4 * ncell, None)
        ginsu.sync(None)

    def stored_bytes(self):
        return 4 * (self.kwords * self.nkey + self.words * (self.nt - self.nkey)) * self.nf

    def _key(self, kind, n):  # keyframe of frame n (the first or second of its segment)
        j, m = divmod(n - 1, self.k)
        i = 2 * j + m
        return self.keys.ptr + 4 * self.kwords * (i * self.nf + kind)

    def _code(self, kind, n):
        j, m = divmod(n - 1, self.k)
        i = n - 1 - 2 * (j + 1)  # residual frames before this one
        return self.codes.ptr + 4 * self.words * (i * self.nf + kind)

    def _slot(self, kind, n):
        m = (n - 1) % self.k
        return self.seg.ptr + 4 * self.ncell * (m * self.nf + kind)

    def _ptr(self, name):
        return prop3d.device_ptr(self.p2, name)

    def _keyframe(self, kind, n, src, stream):
        """code frame n of `src` as a keyframe (src None: only decode it) into the predictor's
        current field, after moving the current one to the old one (second keyframe)"""
        ginsu.sync(stream)  # the block codec works on the null stream
        prop3d.synchronize(self.p2)
        cur = self._ptr(self.cur[kind])
        if (n - 1) % self.k == 1:
            ginsu.copy(self._ptr(self.old[kind]), cur, 4 * self.ncell, None)
        if src is not None and self.cod.wavefd_codec_compress(src, self.n3, self.b3, self.keybits,
                                                               self._key(kind, n)) < 0:
            raise RuntimeError(_lib_error("codec"))
        if self.cod.wavefd_codec_decompress(self._key(kind, n), self.kwords, cur) != 0:
            raise RuntimeError(_lib_error("codec"))
        ginsu.sync(None)

    def _predict(self, n, stream):
        """the prediction of frame n: one step of the predictor, (cur, old) <- (step, cur), then
        cur += w G with the wavelet sample injected after the step that wrote frame n - 1"""
        ginsu.sync(stream)
        lib = self.p2._lib
        if self.nf == 2:
            lib.TimeStepLinear(self.p2.p)
        else:
            lib.TimeStep(self.p2.p)
        self.p2._check()
        prop3d.synchronize(self.p2)
        wn = float(self.wavelet[n - 2])
        for kind in range(self.nf):
            cur = self._ptr(self.cur[kind])
            self.lc.combine(self.ncell, 1.0, cur, cur, wn, self.g[kind].ptr, stream=stream)

    def begin(self, n):
        pass

    def out(self, kind, n):  # where the step writes DP_n
        return self.x[kind].ptr

    def commit(self, kind, n, stream):
        lc, nc = self.lc, self.ncell
        m = (n - 1) % self.k
        if m < 2:
            self._keyframe(kind, n, self.x[kind].ptr, stream)
            return
        if kind == 0:
            self._predict(n, stream)
        pred = self._ptr(self.cur[kind])
        lc.combine(nc, 1.0, self.x[kind].ptr, self.tmp.ptr, -1.0, pred, stream=stream)  # residual
        if self.codec is None:
            lc.encode(self.tmp.ptr, self.shape, self._code(kind, n), stream, self.levels)
            lc.dequantize(self._code(kind, n), self.tmp.ptr, stream, self.shape, self.levels)
        else:
            self.codec.encode(self.tmp.ptr, self.shape, self._code(kind, n), self.scr[0].ptr, self.scr[1].ptr, stream,
                              self.levels)
            self.codec.decode(self._code(kind, n), self.tmp.ptr, self.scr[0].ptr, stream)
        lc.combine(nc, 1.0, pred, pred, 1.0, self.tmp.ptr, stream=stream)  # the reconstruction
        if kind == self.nf - 1:
            ginsu.sync(stream)

    def _decode_segment(self, j, stream):
        n0 = j * self.k + 1
        for n in range(n0, min(n0 + self.k, self.nt + 1)):
            m = n - n0
            if m >= 2:
                self._predict(n, stream)
            for kind in range(self.nf):
                cur = self._ptr(self.cur[kind])
                if m < 2:
                    self._keyframe(kind, n, None, stream)
                else:
                    if self.codec is None:
                        self.lc.dequantize(self._code(kind, n), self.tmp.ptr, stream, self.shape, self.levels)
                    else:
                        self.codec.decode(self._code(kind, n), self.tmp.ptr, self.scr[0].ptr, stream)
                    self.lc.combine(self.ncell, 1.0, cur, cur, 1.0, self.tmp.ptr, stream=stream)
                ginsu.copy(self._slot(kind, n), cur, 4 * self.ncell, stream)
            ginsu.sync(stream)
        self.cached = j

    def get(self, kind, n, stream):
        j = (n - 1) // self.k
        if self.cached != j:
            self._decode_segment(j, stream)
        return self._slot(kind, n)

    def free(self):
        for b in [self.keys, self.codes, self.tmp, self.seg] + self.x + self.g + self.scr:
            b.free()
        prop3d.free(self.p2)


def _lib_error(name):
    from . import _lib
    return _lib.last_error(name)


class BornShot:
    """One shot: source (z, y, x), receiver coordinate arrays, a wavelet of nt samples at dt, all
    in the big grid's coordinates (origin 0)."""

    def __init__(self, src, rcv_z, rcv_y, rcv_x, wavelet):
        self.src = tuple(float(c) for c in src)
        self.rz, self.ry, self.rx = (np.asarray(a, np.float64) for a in (rcv_z, rcv_y, rcv_x))
        self.wavelet = np.ascontiguousarray(np.asarray(wavelet, F32).reshape(1, -1))

    @property
    def nr(self):
        return len(self.rz)


class _ShotRun:
    """The device state of one shot on its window: a propagator in DEVICE mode with the window's
    model, dv, the window image, the DP (DM) snapshot area and the shot's taps."""

    def __init__(self, physics, window, spacing, dt, shot, model, dmodel, nsponge, freesurface, freqQ, qMin,
                 qInterior, snapshots, codec=None, learned=True, checkpoint=None, fused=True, predictive=None,
                 codec_for_residuals=None):
        if checkpoint is not None and (codec is not None or int(checkpoint) < 1):
            raise ValueError("checkpoint: an interval >= 1, and not together with a codec store")
        if predictive is not None and (codec is not None or checkpoint is not None or
                                       int(predictive[0] if isinstance(predictive, tuple) else predictive) < 2):
            raise ValueError("predictive: segments of >= 2 frames, not together with another store")
        self.physics, self.window, self.shot = physics, window, shot
        self.nf = nf = 1 if physics == "iso" else 2
        nz, ny, nx = window.size
        dz, dy, dx = spacing
        self.nt = nt = shot.wavelet.shape[1]
        self.ncell = ncell = window.ncells
        self.nsponge, self.freesurface, self.dt = nsponge, freesurface, dt
        self.ctor = dict(nz=nz, ny=ny, nx=nx, nsponge=nsponge, dz=dz, dy=dy, dx=dx, dt=dt, freesurface=freesurface,
                         freqQ=freqQ, qMin=qMin, qInterior=qInterior)
        self.prop = prop = self.new_prop()
        self.lib = prop._lib
        self.stream = stream = prop._lib.GetStream(prop.p)
        for name in MODEL_FIELDS[physics]:
            ginsu.extract(model[name], window, prop3d.device_ptr(prop, name), stream)
        self.store = None
        self.dv = ginsu.DeviceBuffer(4 * ncell)
        ginsu.extract(dmodel, window, self.dv.ptr, stream)
        self.gimg = ginsu.DeviceBuffer(4 * ncell)
        self.gimg.zero(stream)
        # source and receiver taps in window coordinates
        sz, sy, sx = _shift([shot.src[0]], [shot.src[1]], [shot.src[2]], window, spacing)
        fsi = 1 if freesurface else 0  # 1-based free-surface row; 0: no image terms
        self.src_pts = spacetime.hickscoeffs(F32(dz), F32(dy), F32(dx), 0.0, 0.0, 0.0, nz, ny, nx, sz, sy, sx,
                                             fs_index=fsi)
        rz, ry, rx = _shift(shot.rz, shot.ry, shot.rx, window, spacing)
        self.rcv_pts = spacetime.linearcoeffs(F32(dz), F32(dy), F32(dx), 0.0, 0.0, 0.0, nz, ny, nx, rz, ry, rx,
                                              fs_index=fsi)
        self.born_data = None
        if checkpoint is not None:
            self.store = _CheckpointStore(self, checkpoint, fused)
        elif isinstance(predictive, tuple) and len(predictive) > 1 and predictive[1] == "wave":
            self.store = _WaveStore(self, predictive[0], *predictive[2:], codec=codec_for_residuals)
        elif predictive is not None:
            self.store = _PredictiveStore(window.size, nt, nf, *(predictive if isinstance(predictive, tuple)
                                                                  else (predictive,)),
                                          codec=codec_for_residuals)
        elif codec is None:
            self.store = _ExactStore(ncell, nt, nf, snapshots)
        else:
            self.store = _CodecStore(window.size, nt, nf, codec, learned)

    def new_prop(self):
        prop = _CLS[self.physics](**self.ctor)
        prop3d.coherence(prop, "device")
        return prop

    def begin(self, n):  # forward step n of the background is about to run
        self.store.begin(n)

    def snap(self, kind, n):  # where the step writes DP_n / DM_n (n 1-based)
        return self.store.out(kind, n)

    def dps(self, n):  # where the steps write DP_n (DM_n)
        return tuple(self.store.out(k, n) for k in range(self.nf))

    def commit(self, n):  # after the step that wrote DP_n (DM_n)
        for k in range(self.nf):
            self.store.commit(k, n, self.stream)

    def dps_in(self, n):  # DP_n (DM_n) for reading
        return tuple(self.store.get(k, n, self.stream) for k in range(self.nf))

    def zero_fields(self):
        wf = ("pCur", "pOld") + (("mCur", "mOld") if self.nf == 2 else ())
        for name in wf:
            ginsu.zero(prop3d.device_ptr(self.prop, name), 4 * self.ncell, self.stream)

    def setup_forward(self):
        prop3d.set_sources(self.prop, [self.src_pts], self.shot.wavelet, fields=1 if self.nf == 1 else 3)
        prop3d.set_receivers(self.prop, [], self.nt)

    def setup_born(self):
        self.zero_fields()
        prop3d.set_sources(self.prop, [], np.zeros((1, self.nt), F32))
        prop3d.set_receivers(self.prop, [self.rcv_pts], self.nt)

    def setup_adjoint(self, adjoint_data, keep_born_data):
        """receivers inject the time-reversed data with weights dt^2 v^2 / b (into P only: the
        transpose of the P-field extraction)"""
        prop, nt, dt = self.prop, self.nt, self.dt
        if keep_born_data or adjoint_data is None:
            self.born_data = prop3d.receiver_data(prop, self.shot.nr, nt)
        d = self.born_data if adjoint_data is None else np.ascontiguousarray(adjoint_data, F32)
        prop3d.synchronize(prop)  # the gather below runs on the null stream
        vb = {k: _gather(prop3d.device_ptr(prop, k), self.rcv_pts) for k in ("v", "b")}
        adj_pts = self.rcv_pts.copy()
        adj_pts["c"] = (self.rcv_pts["c"].astype(np.float64) * dt * dt * vb["v"].astype(np.float64) ** 2 /
                        vb["b"].astype(np.float64)).astype(F32)
        self.zero_fields()
        prop3d.set_receivers(prop, [], nt)
        prop3d.set_sources(prop, [adj_pts], np.ascontiguousarray(d[:, ::-1]), fields=1)
        if not keep_born_data:
            self.born_data = None

    def finish(self, image):
        ginsu.accumulate(image, self.window, self.gimg.ptr, self.window.interior(self.nsponge, self.freesurface),
                         self.stream)
        prop3d.synchronize(self.prop)

    def free(self):
        self.dv.free()
        self.gimg.free()
        if self.store is not None:
            self.store.free()
        prop3d.free(self.prop)


def _batch_step(runs, epilogue, lists):
    """one TimeStepBatch launch over the runs' propagators; lists: per-shot array addresses for
    a0, a1, a2 (or fewer)"""
    n = len(runs)
    hs = (ctypes.c_void_p * n)(*[r.prop.p for r in runs])
    arrs = [None, None, None]
    for k, col in enumerate(lists):
        arrs[k] = (ctypes.c_void_p * n)(*col)
    if runs[0].lib.TimeStepBatch(hs, n, epilogue, *arrs) != 0:
        runs[0].prop._check()
        raise RuntimeError("TimeStepBatch failed")


def _run_coupled(runs):
    """passes 1 and 2 of Iso shots as one: TimeStepBornCoupled_V steps the background (saving
    DP_n) and the Born field (with the source of that DP_n) together; the shot's source goes into
    the background, the receivers read the Born field.  Several shots: each step of all of them
    one TimeStepBatch launch (epilogue 4)"""
    nt = runs[0].nt
    for r in runs:
        r.setup_forward()
        prop3d.set_receivers(r.prop, [r.rcv_pts], nt, fields=2)
    for it in range(1, nt + 1):
        if len(runs) > 1:
            for r in runs:
                r.begin(it)
            _batch_step(runs, 4, [[r.dv.ptr for r in runs], [r.snap(0, it) for r in runs]])
            for r in runs:
                r.commit(it)
                r.lib.InjectRegistered(r.prop.p, it)
                r.lib.ExtractRegistered(r.prop.p, it)
            continue
        for r in runs:
            lib, p = r.lib, r.prop.p
            r.begin(it)
            lib.TimeStepBornCoupled_V(p, r.dv.ptr, r.snap(0, it))
            r.commit(it)
            lib.InjectRegistered(p, it)
            lib.ExtractRegistered(p, it)
    for r in runs:
        r.prop._check()


def _run_passes(runs, fused, adjoint_data, keep_born_data, coupled=False):
    """the three passes over a batch of shots of equal nt, in lockstep; with fused and several
    shots, each step of all shots is one launch (TimeStepBatch); coupled (Iso, fused): passes 1
    and 2 as one pass over both wavefields"""
    nt = runs[0].nt
    batched = fused and len(runs) > 1
    if coupled and (not fused or runs[0].nf != 1):
        raise ValueError("the coupled Born pass is for Iso shots with fused steps")
    if coupled:
        _run_coupled(runs)
    # with checkpointed stores, passes 1 and 2 alternate per segment: the background of a segment
    # runs on the store's second propagator, then the Born steps of the segment read its DP_n
    segmented = not coupled and all(isinstance(r.store, _CheckpointStore) for r in runs)
    # 1. background forward, saving DP (and DM)
    for r in ([] if coupled or segmented else runs):
        r.setup_forward()
    for it in range(1, (0 if coupled or segmented else nt) + 1):
        if batched:
            for r in runs:
                r.begin(it)
            _batch_step(runs, 1, list(zip(*[r.dps(it) for r in runs])))
            for r in runs:
                r.commit(it)
                r.lib.InjectRegistered(r.prop.p, it)
            continue
        for r in runs:
            lib, p = r.lib, r.prop.p
            r.begin(it)
            if fused:
                lib.TimeStepSaveDP(p, *r.dps(it))
                r.commit(it)
                lib.InjectRegistered(p, it)
            else:
                lib.TimeStepInjectExtract(p, it)
                lib.ScaleSpatialDerivatives(p)
                ginsu.copy(r.snap(0, it), prop3d.device_ptr(r.prop, "pSpace"), 4 * r.ncell, r.stream)
                if r.nf == 2:
                    ginsu.copy(r.snap(1, it), prop3d.device_ptr(r.prop, "mSpace"), 4 * r.ncell, r.stream)
                r.commit(it)
    for r in runs:
        r.prop._check()
    # 2. Born forward: d = J dv
    for r in ([] if coupled else runs):
        r.setup_born()
    for it in range(1, (0 if coupled else nt) + 1):
        if segmented and (it - 1) % runs[0].store.k == 0:
            for r in runs:
                r.store.background_segment((it - 1) // r.store.k)
        if batched:
            _batch_step(runs, 2, [[r.dv.ptr for r in runs]] + list(zip(*[r.dps_in(it) for r in runs])))
        for r in runs:
            lib, p = r.lib, r.prop.p
            if not batched:
                if fused:
                    lib.TimeStepBornInjection_V(p, r.dv.ptr, *r.dps_in(it))
                else:
                    lib.TimeStep(p)
                    lib.ForwardBornInjection_V(p, r.dv.ptr, *r.dps_in(it))
            lib.ExtractRegistered(p, it)
    for r in runs:
        r.prop._check()
    # 3. adjoint with imaging.  AdjointBornAccumulation_V images pOld after the step, i.e. the
    # step's input field: DP_{nt+2-m} after reversed step m, and one extra step images DP_1.
    for k, r in enumerate(runs):
        r.setup_adjoint(None if adjoint_data is None else adjoint_data[k], keep_born_data)
    for m in range(1, nt + 2):
        n = nt + 2 - m
        if batched:
            if m == 1:
                _batch_step(runs, 0, [])
            else:
                _batch_step(runs, 3, [[r.gimg.ptr for r in runs]] + list(zip(*[r.dps_in(n) for r in runs])))
            if m <= nt:
                for r in runs:
                    r.lib.InjectRegistered(r.prop.p, m)
            continue
        for r in runs:
            lib, p = r.lib, r.prop.p
            if m >= 2 and fused:  # the image reads pOld after the step: the injection below does not touch it
                lib.TimeStepAdjointBornAccumulation_V(p, r.gimg.ptr, *r.dps_in(n))
            else:
                lib.TimeStep(p)
            if m <= nt:
                lib.InjectRegistered(p, m)
            if m >= 2 and not fused:
                lib.AdjointBornAccumulation_V(p, r.gimg.ptr, *r.dps_in(n))
    for r in runs:
        r.prop._check()


def born_shot(physics, window, spacing, dt, shot, model, dmodel, image, nsponge=20, freesurface=True, freqQ=5.0,
              qMin=0.1, qInterior=100.0, snapshots=None, adjoint_data=None, keep_born_data=True, fused=True,
              coupled=None, snapshot_codec=None, learned=True, checkpoint=None, predictive=None,
              residual_codec=None):
    """Run the three passes of one shot on `window` and add its image into `image`.

    model: {'v','b'[,'eps','eta','f']} -> device address of the big grid's field; dmodel: device
    address of the big dv; image: device address of the big gradient (accumulated, +=).
    snapshots: optional DeviceBuffer of at least nt * window cells * 4 B (x2 for VTI), reused
    across shots.  adjoint_data: (nr, nt) float32 data for the adjoint pass (default: the Born
    data of pass 2, i.e. image = J^T J dv).  fused: each pass's per-step pair as one fused step
    (TimeStepSaveDP, TimeStepBornInjection_V, TimeStepAdjointBornAccumulation_V; the same
    results as the separate calls).  coupled: passes 1 and 2 as one pass over both wavefields
    (TimeStepBornCoupled_V; the same results again); by default for Iso with fused steps whose
    snapshots are not stored exactly, since it forms the Born source from the exact DP inside the
    step (a lossy store then affects the imaging only).  With exact storage the two one-wavefield
    passes, on the two-workgroups-per-CU tile, are faster.  snapshot_codec: a
    ``lcodec.LearnedCodec`` to keep DP (DM) as learned-codec codes instead of exactly (decoded
    by its network, or by the plain dequantisation with learned=False).  checkpoint: an interval k
    to recompute DP (DM) from checkpoints of the background every k steps instead of storing them
    (``_CheckpointStore``: the same image bit for bit, ``checkpoint_fields`` of memory instead of
    nt per wavefield).  predictive: a segment length k, or (k, order, keybits), to keep DP (DM) by
    closed-loop temporal prediction (``_PredictiveStore``: a 16-bit keyframe every k frames,
    quantised residuals of a third-order extrapolation in between; lossy, 8.2x fewer stored
    bytes than float32 at k = 48); residual_codec: a ``lcodec.LearnedCodec`` trained on residual
    frames, to code them with the learned decoder.  Returns a dict with the Born
    data, the stored snapshot bytes and timings."""
    if physics not in _CLS:
        raise ValueError("physics must be 'iso' or 'vti'")
    if coupled is None:
        coupled = fused and physics == "iso" and not (checkpoint is None and predictive is None and
                                                      snapshot_codec is None)
    t0 = time.perf_counter()
    run = _ShotRun(physics, window, spacing, dt, shot, model, dmodel, nsponge, freesurface, freqQ, qMin, qInterior,
                   snapshots, snapshot_codec, learned, checkpoint, fused, predictive, residual_codec)
    t1 = time.perf_counter()
    try:
        _run_passes([run], fused, None if adjoint_data is None else [adjoint_data], keep_born_data, coupled)
        run.finish(image)
        t2 = time.perf_counter()
        stored = run.store.stored_bytes() if hasattr(run.store, "stored_bytes") else None
        return {"born_data": run.born_data, "window": window, "t_setup": t1 - t0, "t_passes": t2 - t1,
                "t_total": t2 - t0, "stored_bytes": stored}
    finally:
        run.free()


def born_shots(physics, windows, spacing, dt, shots, model, dmodel, image, nsponge=20, freesurface=True, freqQ=5.0,
               qMin=0.1, qInterior=100.0, batch=4, snapshot_bytes=None, adjoint_data=None, keep_born_data=True,
               fused=True, checkpoint=None, predictive=None):
    """Born modelling and imaging of many shots (one window each, equal nt), several at a time:
    the shots of a batch run their passes together, each time step of all of them one launch
    of the fused kernel (TimeStepBatch).  Batches are formed in order, at most `batch` shots and,
    if snapshot_bytes is given, at most that many bytes of DP (DM) snapshots (a shot larger than
    that runs alone).  Every shot's results equal born_shot's for it; the images are added into
    `image` one shot after the other, in shot order.  checkpoint, predictive: as in born_shot (the
    bytes a shot counts against snapshot_bytes are then its ``checkpoint_fields``, or its
    predictive store's stored bytes and buffers).  Returns the list of
    per-shot Born data (None entries unless keep_born_data)."""
    if physics not in _CLS:
        raise ValueError("physics must be 'iso' or 'vti'")
    if len(windows) != len(shots):
        raise ValueError("one window per shot")
    nf = 1 if physics == "iso" else 2
    nts = {s.wavelet.shape[1] for s in shots}
    if len(nts) > 1:
        raise ValueError("shots of a batch run must share nt")
    nt = nts.pop() if shots else 0
    batch = max(1, min(int(batch), 8))
    groups, cur, cur_bytes = [], [], 0
    for i, w in enumerate(windows):
        if checkpoint is not None:
            b = 4 * w.ncells * checkpoint_fields(nt, nf, checkpoint)
        elif predictive is not None:
            k = predictive[0] if isinstance(predictive, tuple) else predictive
            b = 4 * w.ncells * (nt * nf / 8.0 + (min(k, nt) + 6) * nf)  # codes, decode and encoder buffers
        else:
            b = 4 * w.ncells * nt * nf
        if cur and (len(cur) == batch or (snapshot_bytes is not None and cur_bytes + b > snapshot_bytes)):
            groups.append(cur)
            cur, cur_bytes = [], 0
        cur.append(i)
        cur_bytes += b
    if cur:
        groups.append(cur)
    data = [None] * len(shots)
    # exact storage: one buffer for the largest batch's snapshots, sliced per batch (no allocation
    # per shot)
    need = [(4 * w.ncells * nt * nf + 255) // 256 * 256 for w in windows]
    pool = (ginsu.DeviceBuffer(max(sum(need[i] for i in g) for g in groups))
            if groups and checkpoint is None and predictive is None else None)
    try:
        _run_groups(physics, windows, spacing, dt, shots, model, dmodel, image, nsponge, freesurface, freqQ, qMin,
                    qInterior, adjoint_data, keep_born_data, fused, checkpoint, predictive, groups, need, pool, data)
    finally:
        if pool is not None:
            pool.free()
    return data


def _run_groups(physics, windows, spacing, dt, shots, model, dmodel, image, nsponge, freesurface, freqQ, qMin,
                qInterior, adjoint_data, keep_born_data, fused, checkpoint, predictive, groups, need, pool, data):
    """born_shots' batches, one after the other"""
    for g in groups:
        runs = []
        try:
            off = 0
            for i in g:
                snap = None
                if pool is not None:
                    snap = _Slice(pool.ptr + off, need[i])
                    off += need[i]
                runs.append(_ShotRun(physics, windows[i], spacing, dt, shots[i], model, dmodel, nsponge, freesurface,
                                     freqQ, qMin, qInterior, snap, checkpoint=checkpoint, fused=fused,
                                     predictive=predictive))
            _run_passes(runs, fused, None if adjoint_data is None else [adjoint_data[i] for i in g], keep_born_data,
                        coupled=fused and physics == "iso" and
                        not all(isinstance(r.store, _ExactStore) for r in runs) and
                        (len(runs) == 1 or not any(isinstance(r.store, _CheckpointStore) for r in runs)))
            for i, r in zip(g, runs):
                r.finish(image)
                data[i] = r.born_data
        finally:
            for r in runs:
                r.free()


def model_shot(physics, window, spacing, dt, shot, model, nsponge=20, freesurface=True, freqQ=5.0, qMin=0.1,
               qInterior=100.0):
    """Nonlinear modelling of one shot on `window`: the receivers' P traces (nr, nt)."""
    nz, ny, nx = window.size
    dz, dy, dx = spacing
    nt = shot.wavelet.shape[1]
    nf = 1 if physics == "iso" else 2
    prop = _CLS[physics](nz=nz, ny=ny, nx=nx, nsponge=nsponge, dz=dz, dy=dy, dx=dx, dt=dt, freesurface=freesurface,
                         freqQ=freqQ, qMin=qMin, qInterior=qInterior)
    prop3d.coherence(prop, "device")
    stream = prop._lib.GetStream(prop.p)
    for name in MODEL_FIELDS[physics]:
        ginsu.extract(model[name], window, prop3d.device_ptr(prop, name), stream)
    fsi = 1 if freesurface else 0
    sz, sy, sx = _shift([shot.src[0]], [shot.src[1]], [shot.src[2]], window, spacing)
    rz, ry, rx = _shift(shot.rz, shot.ry, shot.rx, window, spacing)
    src_pts = spacetime.hickscoeffs(F32(dz), F32(dy), F32(dx), 0.0, 0.0, 0.0, nz, ny, nx, sz, sy, sx, fs_index=fsi)
    rcv_pts = spacetime.linearcoeffs(F32(dz), F32(dy), F32(dx), 0.0, 0.0, 0.0, nz, ny, nx, rz, ry, rx,
                                     fs_index=fsi)
    prop3d.set_sources(prop, [src_pts], shot.wavelet, fields=1 if nf == 1 else 3)
    prop3d.set_receivers(prop, [rcv_pts], nt)
    for it in range(1, nt + 1):
        prop._lib.TimeStepInjectExtract(prop.p, it)
    prop._check()
    out = prop3d.receiver_data(prop, shot.nr, nt)
    prop3d.free(prop)
    return out
