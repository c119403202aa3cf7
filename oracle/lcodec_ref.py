"""CPU restatement of the learned snapshot codec -- TEST INFRASTRUCTURE ONLY.

The product (``wavefd_amd/csrc/lcodec.hip``) has no counterpart in the reference (whose
snapshot store is CvxCompress or exact storage, src/compressedio.jl:1-148); this module is what
its kernels are compared against:

* ``quantize`` / ``dequantize``: 8^3 blocks, scale = block max |x| rounded up to bfloat16,
  q = rint(L x / s) clipped to -L..L, x_hat = s (d - L) / L in float32, L = (levels - 1) / 2 for
  levels 11 (the default), 5 or 3 -- bit-exact with the GPU;
* ``forward`` / ``loss_grad``: the decoder network out = q + (F(q) - F(-q)) / 2 with
  F = conv3(relu(conv2(relu(conv1(.))))), 3x3x3 "same" convolutions without bias (zero outside
  the volume), 16 channels; with ``bf16=True`` the
  weights and activations are rounded to bfloat16 where the GPU rounds them, and the sums are
  float64 (the GPU accumulates in float32, so results agree to the accumulation rounding).

Pins (tests/test_lcodec_oracle.py): ``conv`` equals scipy.ndimage.correlate per channel pair;
``loss_grad`` equals central finite differences of ``forward`` (float64, no rounding); the
quantiser's error bound |x - x_hat| <= s / 10 and exactness on the levels.
"""
import numpy as np

CH = 16
NT = 27
LAYER_W = NT * CH * CH
LAYER_P = LAYER_W + CH
NPARAM = 3 * LAYER_P
BS = 8
LV = 5


def bf16(x):
    """round float32 values to bfloat16 (nearest even), returned as float32"""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_up(x):
    """smallest bfloat16 >= x (x >= 0), as float32"""
    x = np.asarray(x, np.float32)
    b = bf16(x)
    u = b.view(np.uint32).copy()
    low = b < x
    u[low] += 1 << 16
    return u.view(np.float32)


DIGITS = {11: 9, 5: 13, 3: 20}  # base-levels digits per 32-bit word: 11^9, 5^13, 3^20 < 2^32


def quantize(field, levels=11):
    """field: (n3, n2, n1) float32 (n1 fastest).  Returns (scales (m3, m2, m1) float32,
    digits (n3, n2, n1) int in 0..levels - 1)"""
    LV = (levels - 1) // 2
    n3, n2, n1 = field.shape
    m3, m2, m1 = (-(-n3 // BS), -(-n2 // BS), -(-n1 // BS))
    pad = np.zeros((m3 * BS, m2 * BS, m1 * BS), np.float32)
    pad[:n3, :n2, :n1] = field
    blocks = pad.reshape(m3, BS, m2, BS, m1, BS)
    mx = np.abs(blocks).max(axis=(1, 3, 5))
    s = bf16_up(mx.astype(np.float32))
    sb = s[:, None, :, None, :, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(sb > 0, np.rint((blocks / sb).astype(np.float32) * np.float32(LV)), 0)
    q = np.clip(q, -LV, LV).astype(np.int64) + LV
    return s, q.reshape(m3 * BS, m2 * BS, m1 * BS)[:n3, :n2, :n1]


def dequantize(scales, digits, levels=11):
    LV = (levels - 1) // 2
    n3, n2, n1 = digits.shape
    s = np.repeat(np.repeat(np.repeat(scales, BS, 0), BS, 1), BS, 2)[:n3, :n2, :n1].astype(np.float32)
    return (s * (digits - LV).astype(np.float32) / np.float32(LV)).astype(np.float32)


def code_words(n1, n2, n3, levels=11):
    """header, bfloat16 scales (2 per word), one choice bit per block, then per block 512 digits
    packed DIGITS[levels] to a word (57, 40 or 26 words)"""
    nb = (-(-n1 // BS)) * (-(-n2 // BS)) * (-(-n3 // BS))
    return 8 + (nb + 1) // 2 + (nb + 31) // 32 + -(-512 // DIGITS[levels]) * nb


def unpack(params):
    """packed parameters -> [(W (27, 16, 16) [t][co][c], b (16,))] * 3"""
    p = np.asarray(params, np.float64)
    out = []
    for l in range(3):
        w = p[l * LAYER_P:l * LAYER_P + LAYER_W].reshape(NT, CH, CH)
        b = p[l * LAYER_P + LAYER_W:(l + 1) * LAYER_P]
        out.append((w, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in layers])


def taps():
    """tap t -> (dx, dy, dz), t = (dx + 1) 9 + (dy + 1) 3 + (dz + 1)"""
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(NT)]


def shift(a, d):
    """a[(n, x, y, z, ...)] -> a shifted so that out[x] = a[x + dx] (zero outside)"""
    out = np.zeros_like(a)
    sl_out, sl_in = [slice(None)], [slice(None)]
    for k, n in zip(d, a.shape[1:4]):
        if k >= 0:
            sl_out.append(slice(0, n - k))
            sl_in.append(slice(k, n))
        else:
            sl_out.append(slice(-k, n))
            sl_in.append(slice(0, n + k))
    out[tuple(sl_out)] = a[tuple(sl_in)]
    return out


def conv(inp, w, b=None):
    """inp (N, X, Y, Z, C); w (27, Co, C); out[v][co] = sum_t sum_c inp[v + off(t)][c] w[t][co][c] + b"""
    out = np.zeros(inp.shape[:4] + (w.shape[1],), np.float64)
    for t, d in enumerate(taps()):
        out += shift(inp, d) @ w[t].T
    if b is not None:
        out += b
    return out


def forward(params, q, bf=True):
    """q: (N, X, Y, Z) float32 volumes.  Returns (out (N, X, Y, Z), cache); the cache holds the
    activations of the 2N volumes q and -q"""
    (w1, b1), (w2, b2), (w3, b3) = unpack(np.asarray(params, np.float64) * mask())
    if bf:  # the GPU's operands: bfloat16 weights and input
        w1, w2, w3 = (bf16(w.astype(np.float32)).astype(np.float64) for w in (w1, w2, w3))
    qq = np.concatenate([q, -np.asarray(q)])
    x0 = np.zeros(qq.shape + (CH,), np.float64)
    x0[..., 0] = bf16(np.asarray(qq, np.float32)) if bf else qq
    a1 = conv(x0, w1, b1)
    h1 = np.maximum(a1, 0)
    if bf:
        h1 = bf16(h1.astype(np.float32)).astype(np.float64)
    a2 = conv(h1, w2, b2)
    h2 = np.maximum(a2, 0)
    if bf:
        h2 = bf16(h2.astype(np.float32)).astype(np.float64)
    a3 = conv(h2, w3, b3)[..., 0]
    n = len(q)
    out = np.asarray(q, np.float64) + 0.5 * (a3[:n] - a3[n:])
    return out, (x0, h1, h2, (w1, w2, w3))


def loss_grad(params, q, target, bf=True):
    """mean squared error and its gradient in the packed parameters"""
    out, (x0, h1, h2, (w1, w2, w3)) = forward(params, q, bf)
    n = out.size
    e = out - np.asarray(target, np.float64)
    loss = float(np.sum(e * e) / n)
    d3 = np.zeros((2 * len(out),) + out.shape[1:] + (CH,))
    d3[:len(out), ..., 0] = 0.5 * 2 * e / n
    d3[len(out):, ..., 0] = -0.5 * 2 * e / n
    if bf:
        d3 = bf16(d3.astype(np.float32)).astype(np.float64)

    def wgrad(dO, inp):
        g = np.zeros((NT, CH, CH))
        for t, d in enumerate(taps()):
            si = shift(inp, d)
            g[t] = np.einsum("nxyzo,nxyzc->oc", dO, si)
        return g, dO.sum(axis=(0, 1, 2, 3))

    def dgrad(dO, w):  # transpose conv: sum_t dO[v - off(t)] w[t]
        out_ = np.zeros(dO.shape[:4] + (w.shape[2],))
        for t, d in enumerate(taps()):
            out_ += shift(dO, tuple(-k for k in d)) @ w[t]
        return out_

    g3, gb3 = wgrad(d3, h2)
    d2 = dgrad(d3, w3) * (h2 > 0)
    if bf:
        d2 = bf16(d2.astype(np.float32)).astype(np.float64)
    g2, gb2 = wgrad(d2, h1)
    d1 = dgrad(d2, w2) * (h1 > 0)
    if bf:
        d1 = bf16(d1.astype(np.float32)).astype(np.float64)
    g1, gb1 = wgrad(d1, x0)
    grad = pack([(g1, gb1), (g2, gb2), (g3, gb3)]) * mask()
    return loss, grad, out


def mask():
    m = np.zeros(NPARAM)
    for l in range(3):
        w = np.ones((NT, CH, CH))
        b = np.zeros(CH)  # no biases
        if l == 0:
            w[:, :, 1:] = 0
        if l == 2:
            w[:, 1:, :] = 0
        m[l * LAYER_P:(l + 1) * LAYER_P] = np.concatenate([w.ravel(), b])
    return m
