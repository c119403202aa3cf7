"""GPU: the Born driver's predictive snapshot store (born.py::_PredictiveStore).

* the combine kernel equals numpy's float32 arithmetic in the same order, bit for bit;
* frames decoded in reverse order (as the adjoint reads them) are the encoder's own
  reconstructions bit for bit, every residual frame is within its quantiser's half step
  (|x - x_hat| <= s / 10 of its residual block) and every keyframe within max|x| / 254;
* on smooth moving frames the residual coding is far more accurate than quantising the
  frames themselves at the same 3.59 bits per value;
* a Born shot with the store gives an image within 1e-3 relative L2 of exact storage, with
  8x fewer stored bytes, also with its residuals decoded by a learned decoder trained on
  residual frames of another model."""
import numpy as np
import pytest

from tests.helpers import make_model, smooth

pytestmark = pytest.mark.gpu

F32 = np.float32


def _dev(a):
    from wavefd_amd import ginsu
    b = ginsu.DeviceBuffer(a.nbytes)
    b.upload(np.ascontiguousarray(a, F32))
    return b


def _get(buf_ptr, n):
    from wavefd_amd import ginsu
    out = np.zeros(n, F32)
    ginsu.copy(out.ctypes.data, buf_ptr, 4 * n)
    ginsu.sync()
    return out


def test_combine_matches_numpy_float32():
    from wavefd_amd import ginsu, lcodec
    rng = np.random.default_rng(0)
    n = 100_003
    a, b, c = (rng.standard_normal(n).astype(F32) for _ in range(3))
    da, db, dc = _dev(a), _dev(b), _dev(c)
    out = ginsu.DeviceBuffer(4 * n)
    for ca, cb, cc, use_b, use_c in ((2.0, -1.0, 0.0, True, False), (1.0, 1.0, 0.0, True, False),
                                     (0.5, -3.0, 1.25, True, True), (1.5, 0.0, 0.0, False, False)):
        lcodec.combine(n, ca, da.ptr, out.ptr, cb, db.ptr if use_b else None, cc, dc.ptr if use_c else None)
        ref = F32(ca) * a
        if use_b:
            ref = ref + F32(cb) * b
        if use_c:
            ref = ref + F32(cc) * c
        assert np.array_equal(_get(out.ptr, n), ref.astype(F32))
    for x in (da, db, dc, out):
        x.free()


def _frames(shape, nt, seed):
    """a smooth pulse moving through the grid, sampled densely in time"""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = []
    for t in range(nt):
        cz, cy, cx = 8 + 0.35 * t, 12 + 0.2 * t, 10 + 0.25 * t
        r2 = ((z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2) / 30.0
        f = (1 - r2) * np.exp(-r2) * np.cos(0.15 * t)
        out.append(np.asfortranarray(f).ravel(order="F").astype(F32))  # nz fastest
    return out


def test_combine_unaligned_slots_match_numpy_float32():
    # segment slots of a frame with an odd cell count start at 4-byte offsets; the combine then
    # takes its element-by-element form, with the same bits
    from wavefd_amd import ginsu, lcodec
    rng = np.random.default_rng(1)
    n = 4 * 2_001 + 1
    a, b, c = (rng.standard_normal(n).astype(F32) for _ in range(3))
    da, db, dc = _dev(np.r_[F32(0), a]), _dev(np.r_[F32(0), F32(0), b]), _dev(np.r_[F32(0), F32(0), F32(0), c])
    out = ginsu.DeviceBuffer(4 * (n + 1))
    for ca, cb, cc in ((2.0, -1.0, 0.0), (0.5, -3.0, 1.25)):
        lcodec.combine(n, ca, da.ptr + 4, out.ptr + 4, cb, db.ptr + 8, cc, dc.ptr + 12)
        ref = (F32(ca) * a + F32(cb) * b + F32(cc) * c).astype(F32)
        assert np.array_equal(_get(out.ptr + 4, n), ref)
    for x in (da, db, dc, out):
        x.free()


@pytest.mark.parametrize("order,learned,shape", [(1, False, (40, 36, 32)), (2, False, (40, 36, 32)),
                                                 (3, False, (40, 36, 32)), (3, True, (40, 36, 32)),
                                                 (3, False, (39, 35, 31))])
def test_round_trip_bounds_and_reverse_decode(order, learned, shape):
    # learned: residuals coded by a (seeded, untrained) learned decoder, whose correction stays
    # inside the quantisation cell, so the same bounds hold.  (39, 35, 31): an odd cell count,
    # whose segment slots are only 4-byte aligned
    from wavefd_amd import _lib, born, ginsu, lcodec
    n = int(np.prod(shape))
    nt, k = 37, 8
    frames = _frames(shape, nt, 1)
    net = lcodec.LearnedCodec(seed=3) if learned else None
    st = born._PredictiveStore(shape, nt, 1, k, order, 8, codec=net)  # 8-bit keyframes: error <= max|x| / 254
    recon, bound = [], []
    hs = ginsu.DeviceBuffer(4 * n)
    for t in range(1, nt + 1):
        ginsu.copy(st.out(0, t), frames[t - 1].ctypes.data, 4 * n)
        st.commit(0, t, None)
        ginsu.sync()
        recon.append(_get(st.recon(0), n))
        if (t - 1) % k:
            assert _lib.load("lcodec").wavefd_lcodec_halfstep(st._code(0, t), hs.ptr, None) == 0
            bound.append(_get(hs.ptr, n))
        else:
            bound.append(np.full(n, np.abs(frames[t - 1]).max() / 254.0, F32))
    for t in range(nt, 0, -1):
        got = _get(st.get(0, t, None), n)
        assert np.array_equal(got, recon[t - 1]), t
        err = np.abs(got.astype(np.float64) - frames[t - 1])
        slack = 1e-6 * np.abs(frames[t - 1]).max()
        assert np.all(err <= bound[t - 1] * (1 + 1e-5) + slack), (t, float((err - bound[t - 1]).max()))
    # against quantising each frame itself at the same rate
    cb = ginsu.DeviceBuffer(4 * lcodec.code_words(shape))
    q = ginsu.DeviceBuffer(4 * n)
    e_plain, e_pred = [], []
    for t in range(1, nt + 1):
        if (t - 1) % k == 0:
            continue
        src = _dev(frames[t - 1])
        lcodec.encode(src.ptr, shape, cb.ptr)
        lcodec.dequantize(cb.ptr, q.ptr)
        nrm = np.linalg.norm(frames[t - 1])
        e_plain.append(np.linalg.norm(_get(q.ptr, n) - frames[t - 1]) / nrm)
        e_pred.append(np.linalg.norm(recon[t - 1] - frames[t - 1]) / nrm)
        src.free()
    print(f"order {order}: frame rel err {np.mean(e_pred):.4f} against {np.mean(e_plain):.4f} quantising the frames",
          flush=True)
    assert np.mean(e_pred) < 0.5 * np.mean(e_plain), (np.mean(e_pred), np.mean(e_plain))
    assert st.stored_bytes() * 6 < 4 * n * nt
    for x in (cb, q, hs):
        x.free()
    st.free()
    if net is not None:
        net.close()


def _born_images(k, order, keybits, shape=(64, 56, 72), nt=240, codec=None):
    """the Born image of one Iso shot with exact snapshots and with the predictive store;
    shape: the big grid (nx, ny, nz), whole-grid window"""
    from wavefd_amd import born, ginsu, spacetime
    nx, ny, nz = shape
    m = make_model("iso", shape, 41)
    dev = {key: _dev(m[key]) for key in born.MODEL_FIELDS["iso"]}
    dv = smooth(np.random.default_rng(42), shape, 40.0, 2).astype(F32)
    dvb = _dev(dv)
    wav = spacetime.causal_ricker(15.0, 1e-3 * np.arange(nt))
    sh = born.BornShot((45.0, 280.0, 320.0), np.full(9, 35.0), np.linspace(150, 420, 9), np.linspace(160, 480, 9), wav)
    win = ginsu.Window((nz, ny, nx), (0, 0, 0), (nz, ny, nx))
    model = {key: b.ptr for key, b in dev.items()}
    imgs, stored = {}, None
    for mode in ("exact", "predictive"):
        img = ginsu.DeviceBuffer(dv.nbytes)
        img.zero()
        r = born.born_shot("iso", win, (10.0, 10.0, 10.0), 1e-3, sh, model, dvb.ptr, img.ptr, nsponge=10,
                           predictive=(k, order, keybits) if mode == "predictive" else None,
                           residual_codec=codec if mode == "predictive" else None)
        imgs[mode] = img.download(np.zeros(shape, F32)).astype(np.float64)
        if mode == "predictive":
            stored = r["stored_bytes"]
        img.free()
    for b in list(dev.values()) + [dvb]:
        b.free()
    rel = np.linalg.norm(imgs["predictive"] - imgs["exact"]) / np.linalg.norm(imgs["exact"])
    ratio = 4 * nx * ny * nz * nt / stored
    print(f"predictive store k={k} order={order} keybits={keybits}: image rel err {rel:.3e}, stored ratio "
          f"{ratio:.2f}", flush=True)
    return rel, ratio


def test_born_image_with_predictive_store_default():
    # the default store (segments of 48, third order, 16-bit keyframes) on a grid of whole
    # 8-cell blocks: within 1e-3 of exact storage with 8x fewer stored bytes
    rel, ratio = _born_images(48, 3, 16)
    assert ratio >= 8.0
    assert rel <= 1e-3


@pytest.mark.parametrize("k,order,keybits", [(32, 3, 8), (32, 2, 16), (64, 3, 16)])
def test_born_image_with_predictive_store_variants(k, order, keybits):
    # reported for comparison; bounded loosely
    rel, ratio = _born_images(k, order, keybits)
    assert ratio >= 7.0 and rel <= 5e-3


def _residual_frames(model_seed, shape, nt, every=3):
    """open-loop third-order temporal residuals DP_n - (3 DP_n-1 - 3 DP_n-2 + DP_n-3) of an Iso
    shot (SaveDP steps), every `every` steps: what the predictive store codes"""
    import wavefd_amd as w
    from wavefd_amd import ginsu, lcodec, spacetime
    nx, ny, nz = shape
    n = nx * ny * nz
    p = w.Prop3DAcoIsoDenQ_DEO2_FDTD(nz=nz, ny=ny, nx=nx, dz=10.0, dy=10.0, dx=10.0, dt=1e-3, nsponge=10,
                                     freesurface=True)
    m = make_model("iso", shape, model_seed)
    w.V(p)[...] = m["v"].T
    w.B(p)[...] = m["b"].T
    w.sync_to_device(p)
    w.coherence(p, "device")
    src = spacetime.hickscoeffs(F32(10), F32(10), F32(10), 0.0, 0.0, 0.0, nz, ny, nx, [45.0], [280.0], [320.0])
    w.set_sources(p, [src], spacetime.causal_ricker(15.0, 1e-3 * np.arange(nt))[None, :])
    ring = [ginsu.DeviceBuffer(4 * n) for _ in range(4)]
    pred = ginsu.DeviceBuffer(4 * n)
    strm = p._lib.GetStream(p.p)
    out = []
    for it in range(1, nt + 1):
        assert p._lib.TimeStepSaveDP(p.p, ring[3].ptr) == 0
        p._lib.InjectRegistered(p.p, it)
        ring = [ring[3]] + ring[:3]
        if it >= 4 and it % every == 0:
            b = ginsu.DeviceBuffer(4 * n)
            lcodec.combine(n, 3.0, ring[1].ptr, pred.ptr, -3.0, ring[2].ptr, 1.0, ring[3].ptr, stream=strm)
            lcodec.combine(n, 1.0, ring[0].ptr, b.ptr, -1.0, pred.ptr, stream=strm)
            out.append(b)
    w.synchronize(p)
    for q in ring + [pred]:
        q.free()
    w.free(p)
    return out


def test_learned_residual_decoder_on_another_model():
    # a decoder trained on residual frames of one model, used on another: the same stored bytes,
    # the image still within 1e-3 of exact storage.  Whether it lowers the error depends on the
    # training frames (measured: 5.1e-4 -> 4.0e-4 in bench.py --codec; 1.25e-4 -> 1.68e-4 here),
    # so the test bounds it rather than asserting a gain
    from wavefd_amd import lcodec
    shape = (64, 56, 72)
    nx, ny, nz = shape
    frames = _residual_frames(7, shape, 240)
    net = lcodec.LearnedCodec(seed=11)
    net.train([b.ptr for b in frames], (nz, ny, nx), iters=600, patch=32, batch=4, lr=1e-3, seed=11)
    for b in frames:
        b.free()
    rel_plain, ratio = _born_images(48, 3, 16)
    rel_learned, ratio_l = _born_images(48, 3, 16, codec=net)
    net.close()
    print(f"learned residual decoder: image rel err {rel_learned:.3e} against {rel_plain:.3e}", flush=True)
    assert ratio_l == ratio
    assert rel_plain <= 1e-3 and rel_learned <= 1e-3


def test_dequantize_with_known_shape_equals_header_read():
    from wavefd_amd import ginsu, lcodec
    shape = (40, 36, 30)
    n = int(np.prod(shape))
    x = smooth(np.random.default_rng(9), shape, 3.0, 2).ravel()
    xb = _dev(x)
    cb = ginsu.DeviceBuffer(4 * lcodec.code_words(shape))
    a, b = ginsu.DeviceBuffer(4 * n), ginsu.DeviceBuffer(4 * n)
    lcodec.encode(xb.ptr, shape, cb.ptr)
    lcodec.dequantize(cb.ptr, a.ptr)
    lcodec.dequantize(cb.ptr, b.ptr, shape=shape)
    assert np.array_equal(_get(a.ptr, n), _get(b.ptr, n))
    for q in (xb, cb, a, b):
        q.free()


@pytest.mark.parametrize("levels", [11, 5])
@pytest.mark.parametrize("kind", ["iso", "vti"])
def test_wave_predicted_store_against_exact(kind, levels):
    """born.py::_WaveStore: frames predicted by one step of a propagator with the window's model
    plus the source's image, residuals quantised (11 or 5 levels); the Born image within 1e-4 of
    exact storage (an error each frame carries is coded away with the next frame's residual), the
    stored bytes those of 2 keyframes and k - 2 residual codes per segment."""
    from wavefd_amd import born, ginsu, lcodec, spacetime
    shape = (64, 56, 60)
    nx, ny, nz = shape
    nt, k = 160, 40
    m = make_model(kind, shape, 51)
    dev = {key: _dev(m[key]) for key in born.MODEL_FIELDS[kind]}
    dv = smooth(np.random.default_rng(52), shape, 40.0, 2).astype(F32)
    dvb = _dev(dv)
    wav = spacetime.causal_ricker(15.0, 1e-3 * np.arange(nt))
    sh = born.BornShot((45.0, 280.0, 300.0), np.full(9, 35.0), np.linspace(150, 420, 9), np.linspace(160, 450, 9), wav)
    win = ginsu.Window((nz, ny, nx), (0, 0, 0), (nz, ny, nx))
    model = {key: b.ptr for key, b in dev.items()}
    imgs, stored = {}, None
    for mode in ("exact", "wave"):
        img = ginsu.DeviceBuffer(dv.nbytes)
        img.zero()
        r = born.born_shot(kind, win, (10.0, 10.0, 10.0), 1e-3, sh, model, dvb.ptr, img.ptr, nsponge=10,
                           predictive=(k, "wave", 16, levels) if mode == "wave" else None)
        imgs[mode] = img.download(np.zeros(shape, F32)).astype(np.float64)
        if mode == "wave":
            stored = r["stored_bytes"]
        img.free()
    for b in list(dev.values()) + [dvb]:
        b.free()
    rel = np.linalg.norm(imgs["wave"] - imgs["exact"]) / np.linalg.norm(imgs["exact"])
    nf = 1 if kind == "iso" else 2
    import ctypes
    from wavefd_amd import _lib
    cod = _lib.load("codec")
    kwords = cod.wavefd_codec_words((ctypes.c_long * 3)(nz, ny, nx), (ctypes.c_long * 3)(8, 8, 8), 16)
    nkey = 2 * (nt // k)
    want = 4 * (kwords * nkey + lcodec.code_words((nz, ny, nx), levels) * (nt - nkey)) * nf
    print(f"wave-predicted store {kind}, {levels} levels: image rel err {rel:.3e}, "
          f"{4 * nx * ny * nz * nt * nf / stored:.2f}x",
          flush=True)
    assert stored == want
    assert np.abs(imgs["exact"]).max() > 0 and rel <= 1e-4
