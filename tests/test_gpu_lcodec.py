"""GPU: the learned snapshot codec (libwavefd_lcodec) against its CPU oracle
(oracle/lcodec_ref.py) and against exact storage.

* the quantiser's code decodes bit for bit like the oracle's, and its norm is the largest scale;
* the matrix-core network's forward pass and its parameter gradient (backward data and weight
  kernels) agree with the oracle's bfloat16-rounded float64 evaluation (the GPU accumulates in
  float32; tolerances in the tests);
* trained on snapshots the propagator generates, the learned decode is closer to the exact
  snapshots than the plain dequantisation on snapshots it did not see;
* the Born image with codec snapshots, judged against exact snapshot storage: both decodes'
  image errors are measured and bounded.
"""
import numpy as np
import pytest

from oracle import lcodec_ref as ref
from tests.helpers import make_model

pytestmark = pytest.mark.gpu

F32 = np.float32


def _dev(a):
    from wavefd_amd import ginsu
    b = ginsu.DeviceBuffer(max(4, a.nbytes))
    b.upload(np.ascontiguousarray(a))
    return b


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("levels", [11, 5, 3])
def test_quantiser_matches_oracle(levels):
    from wavefd_amd import lcodec
    rng = np.random.default_rng(0)
    n3, n2, n1 = 21, 29, 37  # (nx, ny, nz): the last block along every axis is cut
    x = (rng.standard_normal((n3, n2, n1)) * np.exp(rng.uniform(-6, 6, (n3, 1, 1)))).astype(F32)
    x[3:9, 4:12, 5:13] = 0.0  # an all-zero block
    shape = (n1, n2, n3)
    xb = _dev(x)
    from wavefd_amd import ginsu
    cb = ginsu.DeviceBuffer(4 * lcodec.code_words(shape, levels))
    lcodec.encode(xb.ptr, shape, cb.ptr, levels=levels)
    out, out2, hs = (ginsu.DeviceBuffer(x.nbytes) for _ in range(3))
    lcodec.dequantize(cb.ptr, out.ptr)  # levels from the code's header
    lcodec.dequantize(cb.ptr, out2.ptr, shape=shape, levels=levels)
    got = out.download(np.zeros_like(x))
    s, d = ref.quantize(x, levels)
    want = ref.dequantize(s, d, levels)
    assert np.array_equal(got, want)
    assert np.array_equal(out2.download(np.zeros_like(x)), want)
    lib = lcodec._lib.load("lcodec")
    assert lib.wavefd_lcodec_norm(cb.ptr) == float(s.max())
    assert lib.wavefd_lcodec_halfstep(cb.ptr, hs.ptr, None) == 0
    sb = np.repeat(np.repeat(np.repeat(s, 8, 0), 8, 1), 8, 2)[:n3, :n2, :n1]
    assert np.array_equal(hs.download(np.zeros_like(x)), (sb / F32(levels - 1)).astype(F32))
    for b in (xb, cb, out, out2, hs):
        b.free()


def _params(rng, scale=0.15):
    p = rng.standard_normal(ref.NPARAM) * scale
    return (p * ref.mask()).astype(F32)


@pytest.mark.parametrize("norm", [1.0, 3.5])
def test_forward_matches_oracle(norm):
    from wavefd_amd import ginsu, lcodec
    rng = np.random.default_rng(1)
    c = lcodec.LearnedCodec(params=_params(rng))
    vol = (2, 9, 10, 21)
    q = (rng.standard_normal(vol) * norm).astype(F32)
    qb, ob = _dev(q), ginsu.DeviceBuffer(q.nbytes)
    c.forward(qb.ptr, vol, ob.ptr, norm=norm)
    got = ob.download(np.zeros_like(q))
    want, _ = ref.forward(c.params, q / F32(norm))
    want = q + norm * (want - (q / F32(norm)))
    assert _rel(got - q, want - q) <= 5e-3, _rel(got - q, want - q)
    c.close()


def test_gradient_matches_oracle():
    from wavefd_amd import ginsu, lcodec
    rng = np.random.default_rng(2)
    c = lcodec.LearnedCodec(params=_params(rng))
    vol = (2, 8, 11, 19)
    q = rng.standard_normal(vol).astype(F32)
    t = (q + 0.3 * rng.standard_normal(vol)).astype(F32)
    qb, tb, ob = _dev(q), _dev(t), ginsu.DeviceBuffer(q.nbytes)
    loss, g = c.grad(qb.ptr, tb.ptr, vol, ob.ptr, want_grad=True)
    rl, rg, _ = ref.loss_grad(c.params, q, t)
    assert abs(loss - rl) <= 2e-3 * rl, (loss, rl)
    m = ref.mask() > 0
    assert not np.any(g[~m])
    for l in range(3):  # every layer's weights and biases
        sl = slice(l * ref.LAYER_P, (l + 1) * ref.LAYER_P)
        assert _rel(g[sl][m[sl]], rg[sl][m[sl]]) <= 3e-2, (l, _rel(g[sl][m[sl]], rg[sl][m[sl]]))
    # one Adam step lowers the loss on this batch
    c.adam(1e-3)
    l2, _ = c.grad(qb.ptr, tb.ptr, vol, ob.ptr)
    assert l2 < loss
    c.close()


def _make_snaps(seed, shape=(64, 60, 72), nt=300, every=6, start=90):
    """DP snapshots of an Iso shot, made by the propagator's own fused SaveDP steps"""
    import wavefd_amd as w
    from wavefd_amd import ginsu, spacetime as st
    from tests.helpers import host_view
    nx, ny, nz = shape
    p = w.Prop3DAcoIsoDenQ_DEO2_FDTD(nz=nz, ny=ny, nx=nx, dz=12.5, dy=12.5, dx=12.5, dt=1e-3, nsponge=12,
                                     freesurface=True, freqQ=10.0, qMin=0.1, qInterior=100.0)
    m = make_model("iso", shape, seed)
    for k in ("v", "b"):
        host_view(p, k)[:] = m[k].ravel()
    w.coherence(p, "device")
    w.sync_to_device(p)
    src = st.hickscoeffs(F32(12.5), F32(12.5), F32(12.5), 0.0, 0.0, 0.0, nz, ny, nx, [30.0], [ny * 6.25 + 3.1],
                         [nx * 6.25 - 4.2], fs_index=1)
    wav = st.causal_ricker(15.0, 1e-3 * np.arange(nt)).reshape(1, -1).astype(F32)
    w.set_sources(p, [src], wav)
    n = nx * ny * nz
    snaps = []
    dp = ginsu.DeviceBuffer(4 * n)
    for it in range(1, nt + 1):
        assert p._lib.TimeStepSaveDP(p.p, dp.ptr) == 0
        p._lib.InjectRegistered(p.p, it)
        if it >= start and it % every == 0:
            s = ginsu.DeviceBuffer(4 * n)
            ginsu.copy(s.ptr, dp.ptr, 4 * n, p._lib.GetStream(p.p))
            snaps.append(s)
    w.synchronize(p)
    dp.free()
    w.free(p)
    return snaps, (nz, ny, nx)


def _errors(codec, snaps, shape):
    """mean relative errors of the plain and the learned decode of the same codes; checks that
    the learned decode stays in the quantisation cell, |decode - dequantised| <= half a step,
    and is no worse than the plain one in any 8^3 block"""
    from wavefd_amd import ginsu, lcodec
    n = int(np.prod(shape))
    cb = ginsu.DeviceBuffer(4 * lcodec.code_words(shape))
    a, b, scr, hs = (ginsu.DeviceBuffer(4 * n) for _ in range(4))
    lib = lcodec._lib.load("lcodec")
    plain, learned = [], []
    for s in snaps:
        codec.encode(s.ptr, shape, cb.ptr, a.ptr, scr.ptr)  # with the encoder's per-block choice
        lcodec.dequantize(cb.ptr, a.ptr)
        codec.decode(cb.ptr, b.ptr, scr.ptr)
        assert lib.wavefd_lcodec_halfstep(cb.ptr, hs.ptr, None) == 0
        x = s.download(np.zeros(n, F32))
        xa, xb, h = (q.download(np.zeros(n, F32)) for q in (a, b, hs))
        assert np.all(np.abs(xa - x) <= h * (1 + 1e-6))
        assert np.all(np.abs(xb - xa) <= h * (1 + 1e-6))
        nz, ny, nx = shape
        blk = lambda v: np.pad(v.reshape(nx, ny, nz).astype(np.float64) ** 2,  # noqa: E731
                               [(0, -k % 8) for k in (nx, ny, nz)]).reshape(
            -(-nx // 8), 8, -(-ny // 8), 8, -(-nz // 8), 8).sum(axis=(1, 3, 5))
        assert np.all(blk(xb - x) <= blk(xa - x) * (1 + 1e-4) + 1e-30)
        plain.append(_rel(xa, x))
        learned.append(_rel(xb, x))
    for q in (cb, a, b, scr, hs):
        q.free()
    return float(np.mean(plain)), float(np.mean(learned))


def test_training_gains_over_plain_dequantisation():
    from wavefd_amd import lcodec
    train, shape = _make_snaps(31, every=5, start=2)
    test, _ = _make_snaps(32)
    c = lcodec.LearnedCodec(seed=3)
    p0, l0 = _errors(c, test, shape)
    assert abs(p0 - l0) <= 1e-6 * p0  # the untrained decoder is the dequantisation
    losses = c.train([s.ptr for s in train], shape, iters=400, patch=32, batch=4, lr=1e-3, seed=5)
    assert losses[-50:].mean() < 0.8 * losses[:20].mean()
    plain, learned = _errors(c, test, shape)
    print("snapshot rel error: plain %.4f learned %.4f (%d test snapshots)" % (plain, learned, len(test)))
    assert learned < 0.9 * plain
    for s in train + test:
        s.free()
    c.close()


def test_born_image_against_exact_storage():
    from wavefd_amd import born, ginsu, lcodec, spacetime
    train, shape = _make_snaps(41, every=5, start=2)
    c = lcodec.LearnedCodec(seed=4)
    c.train([s.ptr for s in train], shape, iters=500, patch=32, batch=4, lr=1e-3, seed=6)
    for s in train:
        s.free()
    big = (56, 52, 60)  # (nx, ny, nz)
    nx, ny, nz = big
    m = make_model("iso", big, 43)
    dev = {k: _dev(m[k]) for k in ("v", "b")}
    rng = np.random.default_rng(44)
    from tests.helpers import smooth
    dv = smooth(rng, big, 40.0, 2).astype(F32)
    dvb = _dev(dv)
    nt = 220
    shot = born.BornShot((37.5, ny * 6.25, nx * 6.25), np.full(9, 25.0), np.linspace(150, (ny - 12) * 12.5, 9),
                         np.linspace(160, (nx - 12) * 12.5, 9), spacetime.causal_ricker(15.0, 1e-3 * np.arange(nt)))
    win = ginsu.Window((nz, ny, nx), (0, 0, 0), (nz, ny, nx))
    model = {k: b.ptr for k, b in dev.items()}
    imgs = {}
    for mode in ("exact", "plain", "learned"):
        img = ginsu.DeviceBuffer(dv.nbytes)
        img.zero()
        born.born_shot("iso", win, (12.5, 12.5, 12.5), 1e-3, shot, model, dvb.ptr, img.ptr, nsponge=10,
                       snapshot_codec=None if mode == "exact" else c, learned=mode == "learned")
        imgs[mode] = img.download(np.zeros(big, F32))
        img.free()
    ep = _rel(imgs["plain"], imgs["exact"])
    el = _rel(imgs["learned"], imgs["exact"])
    print("Born image rel error vs exact snapshots: plain %.2e learned %.2e (codec %.2fx)" % (
        ep, el, lcodec.ratio((nz, ny, nx))))
    assert np.abs(imgs["exact"]).max() > 0
    # exact storage judges both decodes; the learned decode lowers the snapshot error (tested
    # above) but its residual is a smooth function of the field, which the imaging condition
    # does not average away as it does the plain quantisation's: neither wins on the image
    # everywhere, both stay within these bounds on this shot (measured 0.22 and 0.19)
    assert ep < 0.3 and el < 0.3
    c.close()
