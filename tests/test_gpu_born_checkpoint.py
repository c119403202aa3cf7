"""GPU: the Born driver with DP (DM) recomputed from checkpoints of the background
(born_shot(..., checkpoint=k)) gives the Born data and the image of exact snapshot storage bit for
bit, for Iso (coupled pass, fused pairs, separate calls) and VTI (fused pairs, separate calls),
for intervals that split nt evenly or not, equal 1, equal nt or exceed it, and in batches."""
import numpy as np
import pytest

from tests.helpers import make_model, smooth

pytestmark = pytest.mark.gpu

F32 = np.float32
SP = (10.0, 10.0, 10.0)


def _dev(arr):
    from wavefd_amd import ginsu
    b = ginsu.DeviceBuffer(arr.nbytes)
    b.upload(np.ascontiguousarray(arr, F32))
    return b


def _setup(kind, nt, nshots=1):
    from wavefd_amd import born, ginsu, spacetime
    shape = (56, 50, 60)  # big (nx, ny, nz)
    nx, ny, nz = shape
    m = make_model(kind, shape, 31)
    dev = {k: _dev(m[k]) for k in born.MODEL_FIELDS[kind]}
    dv = smooth(np.random.default_rng(32), shape, 40.0, 2).astype(F32)
    dvb = _dev(dv)
    wav = spacetime.causal_ricker(25.0, 1e-3 * np.arange(nt))
    windows, shots = [], []
    for k, (cy, cx) in enumerate([(200.0, 220.0), (300.0, 330.0), (250.0, 280.0)][:nshots]):
        sh = born.BornShot((45.0, cy, cx), np.full(5, 35.0), np.linspace(cy - 60, cy + 60, 5),
                           np.linspace(cx - 50, cx + 70, 5), wav)
        windows.append(ginsu.window_for_shot((nz, ny, nx), SP, (0.0, 0.0, 0.0), sh.src, sh.rz, sh.ry, sh.rx,
                                             aperture=(40.0 + 20 * k, 30.0), pad=8))
        shots.append(sh)
    return shape, {k: b.ptr for k, b in dev.items()}, dvb, windows, shots, list(dev.values())


def _shot(kind, shape, model, dvb, window, shot, **kw):
    from wavefd_amd import born, ginsu
    img = ginsu.DeviceBuffer(4 * int(np.prod(shape)))
    img.zero()
    data = born.born_shot(kind, window, SP, 1e-3, shot, model, dvb.ptr, img.ptr, nsponge=8, **kw)["born_data"]
    out = img.download(np.zeros(shape, F32))
    img.free()
    return data, out


@pytest.mark.parametrize("kind,form", [("iso", "coupled"), ("iso", "fused"), ("iso", "plain"), ("vti", "fused"),
                                       ("vti", "plain")])
def test_checkpointed_born_equals_exact_storage(kind, form):
    nt = 45
    shape, model, dvb, windows, shots, keep = _setup(kind, nt)
    kw = {"coupled": dict(fused=True, coupled=True), "fused": dict(fused=True, coupled=False),
          "plain": dict(fused=False)}[form]
    d0, g0 = _shot(kind, shape, model, dvb, windows[0], shots[0], **kw)
    assert np.abs(d0).max() > 0 and np.abs(g0).max() > 0
    for k in (1, 7, 16, nt, nt + 5):
        d1, g1 = _shot(kind, shape, model, dvb, windows[0], shots[0], checkpoint=k, **kw)
        assert np.array_equal(d1, d0), (form, k)
        assert np.array_equal(g1, g0), (form, k)
    for b in keep + [dvb]:
        b.free()


@pytest.mark.parametrize("kind", ["iso", "vti"])
def test_predictive_batch_equals_predictive_sequential(kind):
    # the predictive store is lossy, but deterministic: a batch of shots gives each shot the
    # Born data and image it gets alone with the same passes (Iso: the coupled background + Born
    # pass, whose Born source is formed from the exact DP inside the step, in the batch as alone)
    from wavefd_amd import born, ginsu
    nt = 40
    shape, model, dvb, windows, shots, keep = _setup(kind, nt, nshots=3)
    img_s = ginsu.DeviceBuffer(4 * int(np.prod(shape)))
    img_s.zero()
    seq = [born.born_shot(kind, w, SP, 1e-3, s, model, dvb.ptr, img_s.ptr, nsponge=8, predictive=(8, 3, 16))
           ["born_data"] for w, s in zip(windows, shots)]
    img_b = ginsu.DeviceBuffer(4 * int(np.prod(shape)))
    img_b.zero()
    bat = born.born_shots(kind, windows, SP, 1e-3, shots, model, dvb.ptr, img_b.ptr, nsponge=8, batch=3,
                          predictive=(8, 3, 16))
    for x, y in zip(seq, bat):
        assert np.array_equal(x, y)
    gs = img_s.download(np.zeros(shape, F32))
    gb = img_b.download(np.zeros(shape, F32))
    assert np.array_equal(gs, gb) and np.abs(gs).max() > 0
    for b in keep + [dvb, img_s, img_b]:
        b.free()


@pytest.mark.parametrize("kind", ["iso", "vti"])
def test_checkpointed_batch_equals_exact_sequential(kind):
    from wavefd_amd import born, ginsu
    nt = 40
    shape, model, dvb, windows, shots, keep = _setup(kind, nt, nshots=3)
    img_e = ginsu.DeviceBuffer(4 * int(np.prod(shape)))
    img_e.zero()
    seq = [born.born_shot(kind, w, SP, 1e-3, s, model, dvb.ptr, img_e.ptr, nsponge=8)["born_data"]
           for w, s in zip(windows, shots)]
    img_c = ginsu.DeviceBuffer(4 * int(np.prod(shape)))
    img_c.zero()
    bat = born.born_shots(kind, windows, SP, 1e-3, shots, model, dvb.ptr, img_c.ptr, nsponge=8, batch=3,
                          checkpoint=6)
    for x, y in zip(seq, bat):
        assert np.array_equal(x, y)
    ge = img_e.download(np.zeros(shape, F32))
    gc = img_c.download(np.zeros(shape, F32))
    assert np.array_equal(ge, gc) and np.abs(ge).max() > 0
    for b in keep + [dvb, img_e, img_c]:
        b.free()


def test_checkpoint_rejects_bad_interval_and_codec():
    from wavefd_amd import born
    shape, model, dvb, windows, shots, keep = _setup("iso", 10)
    with pytest.raises(ValueError):
        _shot("iso", shape, model, dvb, windows[0], shots[0], checkpoint=0)
    with pytest.raises(ValueError):
        _shot("iso", shape, model, dvb, windows[0], shots[0], checkpoint=3, snapshot_codec=object())
    for b in keep + [dvb]:
        b.free()
    assert born.checkpoint_fields(200, 1, 20) == 2 * 10 + 20 + 6
