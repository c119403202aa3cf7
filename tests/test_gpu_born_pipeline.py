"""GPU: ginsu windows and the Born driver (wavefd_amd/born.py).

* extraction and accumulation are bit-exact copies / single adds at the window offset;
* Born dot test: <J dv, d> = <dv, J^T d> to 1e-5 for Iso and VTI, where J dv is the driver's
  Born data and J^T d its image with the adjoint data d (the driver's loop alignment and
  weights make J^T the exact adjoint of J);
* J is the derivative of the nonlinear modelling: Richardson-extrapolated central differences
  of the modelled traces in v agree with J dv;
* a shot on a window of a larger model gives the same Born data and image, bit for bit, as
  the same shot run on the extracted sub-model.
"""
import numpy as np
import pytest

from tests.helpers import make_model, smooth

pytestmark = pytest.mark.gpu

F32 = np.float32


def _upload(arr):
    from wavefd_amd import ginsu
    buf = ginsu.DeviceBuffer(arr.nbytes)
    buf.upload(np.ascontiguousarray(arr, F32))
    return buf


def _download(buf, shape):
    out = np.zeros(shape, F32)
    return buf.download(out)


def test_extract_accumulate_exact():
    from wavefd_amd import ginsu
    rng = np.random.default_rng(0)
    nx, ny, nz = 23, 29, 37
    big = rng.standard_normal((nx, ny, nz)).astype(F32)  # memory order == Julia (nz, ny, nx)
    w = ginsu.Window((nz, ny, nx), (5, 7, 3), (20, 15, 11))
    bb = _upload(big)
    sub = ginsu.DeviceBuffer(4 * w.ncells)
    ginsu.extract(bb.ptr, w, sub.ptr)
    got = _download(sub, (11, 15, 20))
    assert np.array_equal(got, big[3:14, 7:22, 5:25])
    add = rng.standard_normal((11, 15, 20)).astype(F32)
    ab = _upload(add)
    box = ((2, 17), (1, 15), (4, 9))
    ginsu.accumulate(bb.ptr, w, ab.ptr, box)
    want = big.copy()
    want[3 + 4:3 + 9, 7 + 1:7 + 15, 5 + 2:5 + 17] += add[4:9, 1:15, 2:17]
    assert np.array_equal(_download(bb, big.shape), want)
    # a window reaching past the (true) grid is rejected before any launch
    from wavefd_amd import _lib
    lib = _lib.load("ginsu")
    assert lib.wavefd_ginsu_extract(bb.ptr, nz, ny, nx, 0, 0, 20, sub.ptr, 5, 5, 8, None) == -1
    assert "outside" in _lib.last_error("ginsu")
    assert lib.wavefd_ginsu_accumulate(bb.ptr, nz, ny, nx, 5, 7, 3, ab.ptr, 20, 15, 11, 0, 21, 0, 15, 0, 11,
                                       None) == -1


def _setup(kind, shape, seed):
    """big model (oracle layout (nx, ny, nz)) on the device, a smooth interior dv, a shot"""
    from wavefd_amd import born, spacetime
    nx, ny, nz = shape
    m = make_model(kind, shape, seed)
    dev = {k: _upload(v) for k, v in m.items() if k in born.MODEL_FIELDS[kind]}
    rng = np.random.default_rng(seed + 1)
    dv = smooth(rng, shape, 40.0, 2).astype(F32)
    mask = np.zeros(shape, bool)
    mask[14:nx - 14, 14:ny - 14, 10:nz - 14] = True
    dv[~mask] = 0
    nt = 140
    d = 10.0
    wav = spacetime.causal_ricker(25.0, 1e-3 * np.arange(nt))
    nr = 9
    shot = born.BornShot((45.0, ny * d / 2 + 3.0, nx * d / 2 - 7.0), np.full(nr, 35.0),
                         np.linspace(90, (ny - 10) * d, nr), np.linspace(100, (nx - 10) * d, nr), wav)
    return m, dev, dv, shot


@pytest.mark.parametrize("kind", ["iso", "vti"])
def test_born_dot_product(kind):
    from wavefd_amd import born, ginsu
    shape = (44, 42, 48)
    nx, ny, nz = shape
    m, dev, dv, shot = _setup(kind, shape, 7)
    w = ginsu.Window((nz, ny, nx), (0, 0, 0), (nz, ny, nx))
    dvb = _upload(dv)
    img = ginsu.DeviceBuffer(dv.nbytes)
    img.zero()
    rng = np.random.default_rng(11)
    d2 = rng.standard_normal((shot.nr, shot.wavelet.shape[1])).astype(F32)
    out = born.born_shot(kind, w, (10.0, 10.0, 10.0), 1e-3, shot, {k: b.ptr for k, b in dev.items()}, dvb.ptr,
                         img.ptr, nsponge=8, adjoint_data=d2)
    jdv = out["born_data"].astype(np.float64)
    g = _download(img, shape).astype(np.float64)
    lhs = float(np.sum(jdv * d2))
    rhs = float(np.sum(dv.astype(np.float64) * g))
    assert np.abs(jdv).max() > 0 and np.abs(g).max() > 0
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs)), (lhs, rhs)


@pytest.mark.parametrize("kind", ["iso", "vti"])
def test_born_is_the_linearised_modelling(kind):
    from wavefd_amd import born, ginsu
    shape = (44, 42, 48)
    nx, ny, nz = shape
    m, dev, dv, shot = _setup(kind, shape, 9)
    w = ginsu.Window((nz, ny, nx), (0, 0, 0), (nz, ny, nx))
    sp = (10.0, 10.0, 10.0)
    dvb = _upload(dv)
    img = ginsu.DeviceBuffer(dv.nbytes)
    img.zero()
    jdv = born.born_shot(kind, w, sp, 1e-3, shot, {k: b.ptr for k, b in dev.items()}, dvb.ptr, img.ptr,
                         nsponge=8)["born_data"].astype(np.float64)

    def fwd(h):
        vp = _upload((m["v"] + h * dv).astype(F32))
        model = {k: b.ptr for k, b in dev.items()}
        model["v"] = vp.ptr
        return born.model_shot(kind, w, sp, 1e-3, shot, model, nsponge=8).astype(np.float64)

    fd1 = (fwd(1.0) - fwd(-1.0)) / 2.0
    fd2 = (fwd(0.5) - fwd(-0.5)) / 1.0
    fd = (4 * fd2 - fd1) / 3
    assert np.linalg.norm(jdv - fd) <= 1e-2 * np.linalg.norm(fd), np.linalg.norm(jdv - fd) / np.linalg.norm(fd)


def test_window_equals_direct_run():
    from wavefd_amd import born, ginsu
    shape = (56, 50, 60)  # big (nx, ny, nz)
    nx, ny, nz = shape
    m, dev, dv, shot0 = _setup("iso", shape, 13)
    from wavefd_amd import born as _b
    # a compact spread, so that the window is a real part of the grid
    shot = _b.BornShot((45.0, 240.0, 260.0), np.full(5, 35.0), np.linspace(200, 280, 5), np.linspace(220, 300, 5),
                       shot0.wavelet[0])
    big = (nz, ny, nx)
    w = ginsu.window_for_shot(big, (10.0, 10.0, 10.0), (0.0, 0.0, 0.0), shot.src, shot.rz, shot.ry, shot.rx,
                              aperture=(20.0, 20.0), pad=8)
    assert w.size != big
    dvb = _upload(dv)
    img = ginsu.DeviceBuffer(dv.nbytes)
    img.zero()
    a = born.born_shot("iso", w, (10.0, 10.0, 10.0), 1e-3, shot, {k: b.ptr for k, b in dev.items()}, dvb.ptr,
                       img.ptr, nsponge=8)
    # the same shot on the extracted sub-model, as a grid of its own
    oz, oy, ox = w.origin
    snz, sny, snx = w.size
    sub = {k: _upload(np.ascontiguousarray(m[k][ox:ox + snx, oy:oy + sny, oz:oz + snz])) for k in ("v", "b")}
    sdv = _upload(np.ascontiguousarray(dv[ox:ox + snx, oy:oy + sny, oz:oz + snz]))
    simg = ginsu.DeviceBuffer(4 * w.ncells)
    simg.zero()
    wself = ginsu.Window(w.size, (0, 0, 0), w.size)
    shifted = born.BornShot((shot.src[0] - oz * 10.0, shot.src[1] - oy * 10.0, shot.src[2] - ox * 10.0),
                            shot.rz - oz * 10.0, shot.ry - oy * 10.0, shot.rx - ox * 10.0, shot.wavelet[0])
    b = born.born_shot("iso", wself, (10.0, 10.0, 10.0), 1e-3, shifted, {k: q.ptr for k, q in sub.items()},
                       sdv.ptr, simg.ptr, nsponge=8)
    assert np.array_equal(a["born_data"], b["born_data"])
    gbig = _download(img, shape)
    gsub = _download(simg, (snx, sny, snz))
    (lz0, lz1), (ly0, ly1), (lx0, lx1) = w.interior(8)
    assert np.array_equal(gbig[ox + lx0:ox + lx1, oy + ly0:oy + ly1, oz + lz0:oz + lz1],
                          gsub[lx0:lx1, ly0:ly1, lz0:lz1])
    outside = np.ones(shape, bool)
    outside[ox + lx0:ox + lx1, oy + ly0:oy + ly1, oz + lz0:oz + lz1] = False
    assert not np.any(gbig[outside])
