"""CPU: the learned codec's oracle (oracle/lcodec_ref.py) pinned by textbook routines -- its
convolution is scipy's correlation per channel pair, its gradient is the central finite
difference of its forward pass, its quantiser keeps the stated error bound and is exact on the
levels -- and the library's code size formula (no GPU needed) matches the oracle's and gives at
least 8x."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import lcodec_ref as ref


def test_conv_is_correlation():
    rng = np.random.default_rng(0)
    inp = rng.standard_normal((2, 5, 6, 7, 3))
    w = rng.standard_normal((27, 4, 3))
    b = rng.standard_normal(4)
    got = ref.conv(inp, w, b)
    for n in range(2):
        for co in range(4):
            want = np.full((5, 6, 7), b[co])
            for c in range(3):
                k = np.zeros((3, 3, 3))
                for t, (dx, dy, dz) in enumerate(ref.taps()):
                    k[dx + 1, dy + 1, dz + 1] = w[t, co, c]
                want += ndimage.correlate(inp[n, ..., c], k, mode="constant", cval=0.0)
            assert np.allclose(got[n, ..., co], want, atol=1e-12)


def _params(rng, scale=0.3):
    p = rng.standard_normal(ref.NPARAM) * scale
    return p * ref.mask()


def test_gradient_is_the_derivative():
    rng = np.random.default_rng(1)
    p = _params(rng)
    q = rng.standard_normal((1, 4, 3, 5))
    t = rng.standard_normal((1, 4, 3, 5))
    loss, g, _ = ref.loss_grad(p, q, t, bf=False)
    idx = np.flatnonzero(ref.mask())
    pick = rng.choice(idx, 40, replace=False)
    # the (unused, masked) biases have zero derivative
    pick = np.concatenate([pick, [ref.LAYER_W + 3, ref.LAYER_P + ref.LAYER_W + 5, 2 * ref.LAYER_P + ref.LAYER_W]])
    h = 1e-6
    for i in pick:
        e = np.zeros_like(p)
        e[i] = h
        lp = ref.loss_grad(p + e, q, t, bf=False)[0]
        lm = ref.loss_grad(p - e, q, t, bf=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_decoder_is_odd_and_homogeneous():
    rng = np.random.default_rng(4)
    p = _params(rng)
    q = rng.standard_normal((1, 4, 5, 6))
    f = lambda x: ref.forward(p, x, bf=False)[0] - x  # noqa: E731
    assert np.allclose(f(-q), -f(q), atol=1e-12)
    assert np.allclose(f(2.5 * q), 2.5 * f(q), atol=1e-12)
    assert not np.any(f(np.zeros_like(q)))


def test_padding_carries_no_gradient():
    rng = np.random.default_rng(2)
    p = _params(rng)
    q = rng.standard_normal((1, 3, 3, 4))
    _, g, _ = ref.loss_grad(p, q, q * 1.1, bf=False)
    assert not np.any(g[ref.mask() == 0])


@pytest.mark.parametrize("levels", [11, 5, 3])
def test_quantiser_bound_and_levels(levels):
    half = (levels - 1) // 2
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((19, 13, 21)) * np.exp(rng.uniform(-8, 8, (19, 1, 1)))).astype(np.float32)
    s, d = ref.quantize(x, levels)
    assert d.min() >= 0 and d.max() <= levels - 1
    xh = ref.dequantize(s, d, levels)
    sb = np.repeat(np.repeat(np.repeat(s, 8, 0), 8, 1), 8, 2)[:19, :13, :21]
    assert np.all(np.abs(xh - x) <= sb / (2 * half) * (1 + 1e-6) + 1e-30)
    assert np.all(sb >= np.abs(x))  # scales rounded up: nothing clips
    # values on the levels come back exactly
    sc = ref.bf16(rng.uniform(0.1, 10.0, (2, 2, 3)).astype(np.float32))
    dg = rng.integers(0, levels, (16, 16, 24))
    dg[::8, ::8, ::8] = levels - 1  # every block reaches its scale
    sb = np.repeat(np.repeat(np.repeat(sc, 8, 0), 8, 1), 8, 2)
    lv = (sb * (dg - half).astype(np.float32) / np.float32(half)).astype(np.float32)
    s2, d2 = ref.quantize(lv, levels)
    assert np.array_equal(s2, sc)
    assert np.array_equal(ref.dequantize(s2, d2, levels), lv)


def test_digits_fit_a_word():
    # the packing: DIGITS[levels] base-levels digits must fit 32 bits, one more must not
    for levels, dg in ref.DIGITS.items():
        assert levels ** dg < 2 ** 32 <= levels ** (dg + 1)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0], np.float32)
    got = ref.bf16(x)
    assert np.array_equal(got, np.array([1.0, 1.0, 1.0 + 2 ** -6, -2.5, 0.0], np.float32))
    up = ref.bf16_up(np.array([1.0 + 2 ** -9], np.float32))
    assert up[0] == np.float32(1.0 + 2 ** -7)


@pytest.mark.parametrize("levels", [11, 5, 3])
@pytest.mark.parametrize("shape", [(512, 512, 512), (37, 29, 21), (8, 8, 8)])
def test_code_size(shape, levels):
    from wavefd_amd import lcodec
    nz, ny, nx = shape
    assert lcodec.code_words(shape, levels) == ref.code_words(nz, ny, nx, levels)
    if levels == 11:
        assert lcodec.code_words(shape) == ref.code_words(nz, ny, nx)
    if shape == (512, 512, 512):
        assert lcodec.ratio(shape, levels) >= {11: 8.0, 5: 12.5, 3: 19.0}[levels]


def test_code_size_rejects_other_levels():
    from wavefd_amd import lcodec
    for levels in (2, 4, 7, 13):
        with pytest.raises(ValueError):
            lcodec.code_words((16, 16, 16), levels)
