// libwavefd_lcodec -- learned snapshot codec for the Born/RTM gradient's forward snapshots.
//
// The reference stores snapshots exactly or through CvxCompress (src/compressedio.jl:1-148;
// CvxCompress is un-vendored).  This codec has two parts:
//   * encoder (fixed): 8^3 blocks (cut at the grid's ends), one scale per block -- the block's
//     max |x| rounded up to bfloat16 -- and 11 levels per value, q = rint(5 x / s) in -5..5,
//     packed 9 base-11 digits per 32-bit word (11^9 < 2^32): 57 words per block, i.e. 3.59 bits
//     per value with the scale and the block's choice bit, 8.9x smaller than float32;
//   * decoder (learned): the dequantised field q_hat = s (d - 5) / 5 is corrected by a small
//     3D convolutional network, out = q_hat + g f(q_hat / g) (g: the largest block scale, kept
//     in the code's header), f(q) = (F(q) - F(-q)) / 2 with F = conv3 . relu . conv2 . relu .
//     conv1, each conv 3x3x3 without bias, 16 channels (conv1 reads one channel, conv3 writes
//     one), trained on snapshots the propagator generates (MSE against the exact snapshot).
//     f is odd and positively homogeneous, like the wave equation's response to its source and
//     like the quantiser: f(0) = 0 and f(-q) = -f(q), so the decoder adds no offset and no
//     rectified part, errors that would add up coherently in the imaging condition.  The
//     correction is clamped to half a quantisation step, |out - q_hat| <= s / 10: the exact
//     value lies in that cell, so the decode never leaves it (a consistent reconstruction).
//   * encoder's choice: the encoder, which knows the exact field, runs the decoder and keeps
//     one bit per block saying whether the learned or the plain decode is closer there; the
//     learned decode of a block is then never worse than the plain one (the network, trained on
//     typical blocks, would flatten the sharp peaks of the source region).  1 bit per 512
//     values.
// The network's convolutions, forward and backward (data and weights), run on the matrix
// cores: gfx950's v_mfma_f32_16x16x32_bf16 with bfloat16 weights and activations and float32
// accumulation (two taps of 16 channels, or 32 voxels of the weight gradient, per instruction;
// the input tiles staged in LDS).  Every conv is computed as 16 -> 16 channels with zero-padded
// weights, so one kernel serves all three layers and their transposes.
//
// Layouts: a volume batch is npatch volumes of X x Y x Z voxels, z fastest (the propagators'
// x-slowest layout with X = nx, Y = ny, Z = nz).  Activations and activation gradients are
// bfloat16, 16 channels per voxel, channel fastest.  Weights W[t][co][c] (t = tap, 27), with
// the matrix-core B operand layout Wb[t][n][k] (n output, k input channel).
//
// Matrix-core operand layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l supplies
// A[m = l % 16][k = 8 (l / 16) + j] and B[k = 8 (l / 16) + j][n = l % 16] for j = 0..7, and
// holds D[m = 4 (l / 16) + j][n = l % 16], j = 0..3.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

std::string g_err;

#define LC_TRY(x)                                                        \
    do {                                                                 \
        hipError_t e_ = (x);                                             \
        if (e_ != hipSuccess) {                                          \
            g_err = std::string(#x) + ": " + hipGetErrorString(e_);      \
            return -1;                                                   \
        }                                                                \
    } while (0)

constexpr uint32_t MAGIC = 0x5746444cu;  // "WFDL"
constexpr int BS = 8, BV = BS * BS * BS;  // block extent and values per block
constexpr int LV = 5;                     // levels -LV..LV of the default quantiser (11 levels)
constexpr int CH = 16;                    // channels
constexpr int NT = 27;                    // taps
constexpr int LAYER_W = NT * CH * CH;     // weights per layer (padded)
constexpr int LAYER_P = LAYER_W + CH;     // weights + bias per layer
constexpr int NPARAM = 3 * LAYER_P;

// ---- bfloat16 ---------------------------------------------------------------------------
__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
__host__ __device__ inline uint16_t f2bf_up(float f) {  // smallest bfloat16 >= f (f >= 0, finite)
    uint16_t b = f2bf(f);
    if (bf2f(b) < f) b = (uint16_t)(b + 1);
    return b;
}

// ---- quantiser --------------------------------------------------------------------------
// 2 lv + 1 levels for lv = 5 (11 levels, the default), 2 (5 levels) or 1 (3 levels): q_digits
// base-(2 lv + 1) digits per 32-bit word (11^9, 5^13, 3^20 < 2^32), q_words words per 8^3 block
// (57, 40, 26: 3.56, 2.5 and 1.63 bits per value)
__host__ __device__ constexpr int q_digits(int lv) { return lv == 5 ? 9 : (lv == 2 ? 13 : 20); }
__host__ __device__ constexpr int q_words(int lv) { return (BV + q_digits(lv) - 1) / q_digits(lv); }
inline bool q_levels_ok(long lv) { return lv == 5 || lv == 2 || lv == 1; }

struct QGeo {
    long n1, n2, n3, m1, m2, m3;
    int lv;
    __host__ __device__ long nblocks() const { return m1 * m2 * m3; }
};
inline QGeo qgeo(long n1, long n2, long n3, int lv = LV) {
    return {n1, n2, n3, (n1 + BS - 1) / BS, (n2 + BS - 1) / BS, (n3 + BS - 1) / BS, lv};
}
__host__ __device__ inline long scale_words(const QGeo &g) { return (g.nblocks() + 1) / 2; }
// one bit per block after the scales: 1 = the block's decode is the learned one
__host__ __device__ inline long flag_off(const QGeo &g) { return 8 + scale_words(g); }
__host__ __device__ inline long pay_off(const QGeo &g) { return flag_off(g) + (g.nblocks() + 31) / 32; }
inline long code_words(const QGeo &g) { return pay_off(g) + (long)q_words(g.lv) * g.nblocks(); }

__device__ __forceinline__ long block_elem(const QGeo &g, long blk, int e, bool &in) {
    const long k1 = blk % g.m1, k2 = (blk / g.m1) % g.m2, k3 = blk / (g.m1 * g.m2);
    const long i1 = k1 * BS + (e & 7), i2 = k2 * BS + ((e >> 3) & 7), i3 = k3 * BS + (e >> 6);
    in = i1 < g.n1 && i2 < g.n2 && i3 < g.n3;
    return in ? (i3 * g.n2 + i2) * g.n1 + i1 : 0;
}

// Four blocks along n1 per workgroup, the layout of k_dequantize: lane tid reads the 8 values of
// block tid % 4 in row tid / 4 (four lanes, one 128-byte line); each block's maximum is a shuffle
// over the 16 lanes of a wave that hold it, then over the 4 waves; one wave packs each block.
template <int L>
__global__ void __launch_bounds__(256) k_encode(QGeo g, const float *__restrict__ x, uint32_t *__restrict__ code) {
    constexpr int D = q_digits(L), W = q_words(L);
    const long m1q = (g.m1 + 3) / 4;
    const long wg = blockIdx.x;
    const long k1q = wg % m1q, rest = wg / m1q, k2 = rest % g.m2, k3 = rest / g.m2;
    const int tid = threadIdx.x, w = tid >> 6, t = tid & 63;
    const int q = tid & 3, r = tid >> 2;
    __shared__ uint8_t dig[4][BV];
    __shared__ float red[4][4];
    __shared__ float sc[4];
    const long k1 = k1q * 4 + q;
    const long i2 = k2 * BS + (r & 7), i3 = k3 * BS + (r >> 3);
    const bool row_in = k1 < g.m1 && i2 < g.n2 && i3 < g.n3;
    const long base = row_in ? (i3 * g.n2 + i2) * g.n1 + k1 * BS : 0;
    const long nrow = g.n1 - k1 * BS;
    float v[BS];
    float m = 0.0f;
#pragma unroll
    for (int j = 0; j < BS; j++) {
        v[j] = (row_in && j < nrow) ? x[base + j] : 0.0f;
        m = fmaxf(m, fabsf(v[j]));
    }
#pragma unroll
    for (int off = 4; off < 64; off <<= 1) m = fmaxf(m, __shfl_xor(m, off));
    if (t < 4) red[w][t] = m;  // lane t < 4 holds block t's maximum over this wave
    __syncthreads();
    if (tid < 4) {
        const float mm = fmaxf(fmaxf(red[0][tid], red[1][tid]), fmaxf(red[2][tid], red[3][tid]));
        sc[tid] = bf2f(f2bf_up(mm));
    }
    __syncthreads();
    const float s = sc[q];
#pragma unroll
    for (int j = 0; j < BS; j++) {
        int qv = 0;
        if (s > 0.0f) {
            qv = (int)rintf(v[j] / s * (float)L);
            qv = qv < -L ? -L : (qv > L ? L : qv);
        }
        dig[q][r * BS + j] = (uint8_t)(qv + L);
    }
    __syncthreads();
    const long k1w = k1q * 4 + w;  // the block wave w packs
    if (k1w >= g.m1) return;
    const long blk = (k3 * g.m2 + k2) * g.m1 + k1w;
    if (t == 0) reinterpret_cast<uint16_t *>(code + 8)[blk] = f2bf_up(sc[w]);  // the largest goes to word 7 (k_norm)
    if (t < W) {
        uint32_t wd = 0, p = 1;
        for (int j = 0; j < D; j++) {
            const int e = t * D + j;
            wd += (e < BV ? (uint32_t)dig[w][e] : (uint32_t)L) * p;
            p *= (uint32_t)(2 * L + 1);
        }
        code[pay_off(g) + (long)W * blk + t] = wd;
    }
}

template <int L>
__device__ __forceinline__ float dequant(float s, int d) { return s * (float)(d - L) / (float)L; }

// half a quantisation step of every value's block, s / (2 lv) (the bound of |x - x_hat|)
__global__ void __launch_bounds__(256) k_halfstep(QGeo g, const uint32_t *__restrict__ code, float *__restrict__ out) {
    const long blk = blockIdx.x;
    const uint16_t *scales = reinterpret_cast<const uint16_t *>(code + 8);
    const float h = bf2f(scales[blk]) / (float)(2 * g.lv);
    for (int e = threadIdx.x; e < BV; e += blockDim.x) {
        bool in;
        const long k = block_elem(g, blk, e, in);
        if (in) out[k] = h;
    }
}

// the predictive store's encoder: pred = ca a + cb b + cc c (b, c optional, in that order) and
// res = x - pred in one pass (the values k_lincomb would give for each)
__global__ void k_pred_res(long n, float ca, const float *__restrict__ a, float cb, const float *__restrict__ b,
                           float cc, const float *__restrict__ c, const float *__restrict__ x,
                           float *__restrict__ pred, float *__restrict__ res) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float v = ca * a[i];
        if (b) v = v + cb * b[i];
        if (c) v = v + cc * c[i];
        pred[i] = v;
        res[i] = x[i] + -1.0f * v;
    }
}

// header word 7: the largest block scale, as the bit pattern of a float (non-negative bfloat16
// scales order as their bits); one workgroup, after k_encode (a single atomic word updated by
// every block serialises them)
__global__ void __launch_bounds__(1024) k_norm(QGeo g, uint32_t *__restrict__ code) {
    __shared__ uint32_t red[1024];
    const uint16_t *scales = reinterpret_cast<const uint16_t *>(code + 8);
    uint32_t m = 0;
    for (long b = threadIdx.x; b < g.nblocks(); b += 1024) m = max(m, (uint32_t)scales[b]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) code[7] = red[0] << 16;
}

// Four blocks along n1 per workgroup (one wave unpacks each block's 57 words into LDS), then the
// stores go row by row: lane tid writes the 8 values of block tid % 4 in row tid / 4, so four
// consecutive lanes write 32 consecutive floats of the grid (one 128-byte line) instead of one
// 32-byte piece per lane.
template <int L>
__global__ void __launch_bounds__(256) k_dequantize(QGeo g, const uint32_t *__restrict__ code, float *__restrict__ out) {
    constexpr int D = q_digits(L), W = q_words(L);
    const long m1q = (g.m1 + 3) / 4;
    const long wg = blockIdx.x;
    const long k1q = wg % m1q, rest = wg / m1q, k2 = rest % g.m2, k3 = rest / g.m2;
    const int tid = threadIdx.x, w = tid >> 6, t = tid & 63;
    __shared__ uint8_t dig[4][D * W];
    __shared__ float sc[4];
    const uint16_t *scales = reinterpret_cast<const uint16_t *>(code + 8);
    const long k1w = k1q * 4 + w;  // the block wave w unpacks
    if (k1w < g.m1) {
        const long blk = (k3 * g.m2 + k2) * g.m1 + k1w;
        if (t == 0) sc[w] = bf2f(scales[blk]);
        if (t < W) {
            uint32_t word = code[pay_off(g) + (long)W * blk + t];
            for (int j = 0; j < D; j++) {
                dig[w][t * D + j] = (uint8_t)(word % (uint32_t)(2 * L + 1));
                word /= (uint32_t)(2 * L + 1);
            }
        }
    }
    __syncthreads();
    const int q = tid & 3, r = tid >> 2;  // block k1q * 4 + q, row r of it
    const long k1 = k1q * 4 + q;
    const long i2 = k2 * BS + (r & 7), i3 = k3 * BS + (r >> 3);
    if (k1 >= g.m1 || i2 >= g.n2 || i3 >= g.n3) return;
    const long base = (i3 * g.n2 + i2) * g.n1 + k1 * BS;
    const long nrow = g.n1 - k1 * BS;
    const float s = sc[q];
#pragma unroll
    for (int j = 0; j < BS; j++)
        if (j < nrow) out[base + j] = dequant<L>(s, dig[q][r * BS + j]);
}

static inline unsigned dequantize_grid(const QGeo &g) { return (unsigned)(((g.m1 + 3) / 4) * g.m2 * g.m3); }

static inline void launch_dequantize(const QGeo &g, const uint32_t *code, float *out, hipStream_t s) {
    if (g.lv == 1)
        k_dequantize<1><<<dequantize_grid(g), 256, 0, s>>>(g, code, out);
    else if (g.lv == 2)
        k_dequantize<2><<<dequantize_grid(g), 256, 0, s>>>(g, code, out);
    else
        k_dequantize<5><<<dequantize_grid(g), 256, 0, s>>>(g, code, out);
}

// encoder's choice per block: flag the block when the learned decode o is closer to the exact x
// than the plain dequantisation qh (sums of squares over the block's values)
__global__ void __launch_bounds__(256) k_select(QGeo g, const float *__restrict__ x, const float *__restrict__ qh,
                                                const float *__restrict__ o, uint32_t *__restrict__ code) {
    const long blk = blockIdx.x;
    __shared__ float rq[256], ro[256];
    float eq = 0.0f, eo = 0.0f;
    for (int e = threadIdx.x; e < BV; e += 256) {
        bool in;
        const long k = block_elem(g, blk, e, in);
        if (in) {
            const float a = qh[k] - x[k], b = o[k] - x[k];
            eq += a * a;
            eo += b * b;
        }
    }
    rq[threadIdx.x] = eq;
    ro[threadIdx.x] = eo;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            rq[threadIdx.x] += rq[threadIdx.x + st];
            ro[threadIdx.x] += ro[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && ro[0] < rq[0]) atomicOr(code + flag_off(g) + (blk >> 5), 1u << (blk & 31));
}

// decoder's side: blocks without the flag take the plain dequantisation
__global__ void __launch_bounds__(256) k_apply_flags(QGeo g, const uint32_t *__restrict__ code,
                                                     const float *__restrict__ qh, float *__restrict__ out) {
    const long blk = blockIdx.x;
    if ((code[flag_off(g) + (blk >> 5)] >> (blk & 31)) & 1u) return;
    for (int e = threadIdx.x; e < BV; e += 256) {
        bool in;
        const long k = block_elem(g, blk, e, in);
        if (in) out[k] = qh[k];
    }
}

// ---- convolutions on the matrix cores --------------------------------------------------
typedef float f4 __attribute__((ext_vector_type(4)));

struct Vol {
    long npatch, X, Y, Z;
    __host__ __device__ long nvox() const { return npatch * X * Y * Z; }
};

// voxel (p, x, y, z) shifted by tap t (dx, dy, dz in -1..1); -1 outside the volume
__device__ __forceinline__ long shifted(const Vol &v, long p, long x, long y, long z, int t) {
    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    const long xx = x + dx, yy = y + dy, zz = z + dz;
    if (xx < 0 || xx >= v.X || yy < 0 || yy >= v.Y || zz < 0 || zz >= v.Z) return -1;
    return ((p * v.X + xx) * v.Y + yy) * v.Z + zz;
}


enum ConvOut { OUT_RELU = 0, OUT_RAW = 1, OUT_MASK = 2 };

// out[v][n] = sum_t sum_k in[v + t][k] Wb[t][n][k] (+ epilogue), on the matrix cores of gfx950:
// v_mfma_f32_16x16x32_bf16 takes two taps (2 x 16 input channels) per instruction, 14 per 16-voxel
// output tile (the 28th tap zero).  A workgroup (4 waves) owns one x-plane, CONV_BY y-rows and 16 z
// of one volume: it stages the 3 x (CONV_BY + 2) x 18 input voxels around them in LDS (zero outside
// the volume), each wave keeps the layer's 14 weight operands in registers and computes CONV_BY / 4
// output tiles, its A operands read from LDS (16 bytes per lane and instruction).
//   OUT_RELU:     act[v][n] = bf16(relu(acc + bias[n]))
//   OUT_RAW:      res[v] = acc + bias[0], channel 0 only, float32
//   OUT_MASK:     act[v][n] = bf16(acc) where mask[v][n] > 0, else 0 (relu backward)
// Operand layout (wave64): lane l supplies A[m = l % 16][k = 8 (l / 16) + j] and
// B[k = 8 (l / 16) + j][n = l % 16], j = 0..7, and holds D[m = 4 (l / 16) + j][n = l % 16], j = 0..3;
// k < 16 is the pair's first tap (channel k), k >= 16 its second (channel k - 16).
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
constexpr int CONV_BY = 8, CONV_SY = CONV_BY + 2, CONV_SZ = 18;
constexpr int CONV_STAGE = 3 * CONV_SY * CONV_SZ * 2;  // 16-byte halves (8 channels) of the staged voxels

__host__ __device__ inline long conv_blocks(const Vol &v) {
    return v.npatch * v.X * ((v.Y + CONV_BY - 1) / CONV_BY) * ((v.Z + 15) / 16);
}

template <int MODE>
__global__ void __launch_bounds__(256) k_conv(Vol vol, const uint16_t *__restrict__ in, const uint16_t *__restrict__ wb,
                                              const float *__restrict__ bias, const uint16_t *__restrict__ mask,
                                              uint16_t *__restrict__ act, float *__restrict__ res) {
    __shared__ uint4 st[CONV_STAGE];
    const long nzt = (vol.Z + 15) / 16, nyt = (vol.Y + CONV_BY - 1) / CONV_BY;
    long b = blockIdx.x;
    const long zt = b % nzt;
    b /= nzt;
    const long ytl = b % nyt;
    b /= nyt;
    const long x = b % vol.X, p = b / vol.X;
    const long z0 = zt * 16, y0 = ytl * CONV_BY;
    for (int i = threadIdx.x; i < CONV_STAGE; i += 256) {
        const int h = i & 1, v = i >> 1;
        const int zi = v % CONV_SZ, yi = (v / CONV_SZ) % CONV_SY, xi = v / (CONV_SZ * CONV_SY);
        const long xx = x + xi - 1, yy = y0 + yi - 1, zz = z0 + zi - 1;
        uint4 val = make_uint4(0u, 0u, 0u, 0u);
        if (xx >= 0 && xx < vol.X && yy >= 0 && yy < vol.Y && zz >= 0 && zz < vol.Z)
            val = *reinterpret_cast<const uint4 *>(in + (((p * vol.X + xx) * vol.Y + yy) * vol.Z + zz) * CH + 8 * h);
        st[i] = val;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const int hsel = g & 1, tsel = g >> 1;  // channel half, tap of the pair
    bf8 wreg[14];
#pragma unroll
    for (int q = 0; q < 14; q++) {
        const int t = 2 * q + tsel;
        const uint4 u = t < NT ? *reinterpret_cast<const uint4 *>(wb + (t * CH + r) * CH + 8 * hsel)
                               : make_uint4(0u, 0u, 0u, 0u);
        wreg[q] = __builtin_bit_cast(bf8, u);
    }
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < CONV_BY / 4; rr++) {
        const int yl = w * (CONV_BY / 4) + rr;  // this wave's y-row in the block (uniform)
        const long y = y0 + yl;
        if (y >= vol.Y) break;
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int q = 0; q < 14; q++) {
            const int t = 2 * q + tsel;
            const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
            const int vi = ((1 + dx) * CONV_SY + (yl + 1 + dy)) * CONV_SZ + (r + 1 + dz);
            const uint4 u = t < NT ? st[2 * vi + hsel] : make_uint4(0u, 0u, 0u, 0u);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, u), wreg[q], acc, 0, 0, 0);
        }
        const int n = r;  // D column
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long z = z0 + 4 * g + j;
            if (z >= vol.Z) continue;
            const long v = ((p * vol.X + x) * vol.Y + y) * vol.Z + z;
            const float a = acc[j];
            if constexpr (MODE == OUT_RELU) {
                const float hv = a + bias[n];
                act[v * CH + n] = f2bf(hv > 0.0f ? hv : 0.0f);
            } else if constexpr (MODE == OUT_RAW) {
                if (n == 0) res[v] = a + bias[0];
            } else {
                act[v * CH + n] = bf2f(mask[v * CH + n]) > 0.0f ? f2bf(a) : (uint16_t)0;
            }
        }
    }
}

// partial weight gradients: G[t][co][c] = sum_v dO[v][co] in[v + t][c] over the 32-voxel z-tiles of
// one workgroup (blockIdx.x) for tap blockIdx.y, v_mfma_f32_16x16x32_bf16 with K = 32 voxels (lane l
// supplies A[co = l % 16][voxel 8 (l / 16) + j] and B[voxel 8 (l / 16) + j][c = l % 16]); each wave
// takes every 4th tile of the range and writes its own partial, part[(blk * 4 + wave) * 27 + t][co][c]
__host__ __device__ inline long wgrad_tiles(const Vol &v) { return v.npatch * v.X * v.Y * ((v.Z + 31) / 32); }

__global__ void __launch_bounds__(256) k_wgrad(Vol vol, const uint16_t *__restrict__ dO, const uint16_t *__restrict__ in,
                                               long tiles_per_blk, float *__restrict__ part) {
    const int t = blockIdx.y, wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const long zt = (vol.Z + 31) / 32, nt = wgrad_tiles(vol);
    const long t0 = (long)blockIdx.x * tiles_per_blk;
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (long tile = t0 + wave; tile < t0 + tiles_per_blk && tile < nt; tile += 4) {
        const long row = tile / zt, z0 = (tile - row * zt) * 32;
        const long y = row % vol.Y, x = (row / vol.Y) % vol.X, p = row / (vol.Y * vol.X);
        uint16_t av[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const long z = z0 + 8 * g + j;  // K index = voxel
            av[j] = 0;
            bv[j] = 0;
            if (z < vol.Z) {
                const long v = ((p * vol.X + x) * vol.Y + y) * vol.Z + z;
                av[j] = dO[v * CH + r];  // A[m = co = r][k = voxel]
                const long sh = shifted(vol, p, x, y, z, t);
                if (sh >= 0) bv[j] = in[sh * CH + r];  // B[k = voxel][n = c = r]
            }
        }
        bf8 a, b;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            a[j] = __builtin_bit_cast(__bf16, av[j]);
            b[j] = __builtin_bit_cast(__bf16, bv[j]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    float *o = part + ((long)(blockIdx.x * 4 + wave) * NT + t) * CH * CH;
    for (int j = 0; j < 4; j++) o[(4 * g + j) * CH + r] = acc[j];  // D[m = co][n = c]
}

// partial bias gradients: part[blk][co] = sum_v dO[v][co] over the workgroup's voxel range
__global__ void __launch_bounds__(256) k_bgrad(long nvox, const uint16_t *__restrict__ dO, long vox_per_blk,
                                               float *__restrict__ part) {
    __shared__ float red[256];
    const int co = threadIdx.x & 15, sub = threadIdx.x >> 4;  // 16 sub-lanes per channel
    const long v0 = (long)blockIdx.x * vox_per_blk;
    float s = 0.0f;
    for (long v = v0 + sub; v < v0 + vox_per_blk && v < nvox; v += 16) s += bf2f(dO[v * CH + co]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < 16) {
        float a = 0.0f;
        for (int k = 0; k < 16; k++) a += red[k * 16 + threadIdx.x];
        part[blockIdx.x * CH + threadIdx.x] = a;
    }
}

// sum of nparts partial vectors of length len, in part order: out[i] = sum_p part[p][i]
__global__ void k_reduce_parts(const float *__restrict__ part, long nparts, long len, float *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    float s = 0.0f;
    for (long p = 0; p < nparts; p++) s += part[p * len + i];
    out[i] = s;
}

// one-channel float32 field, times inv_norm -> 16-channel bfloat16 activations (channel 0) of
// the volumes 0..n-1, and of its negation for the volumes n..2n-1
__global__ void k_pack1(long n, const float *__restrict__ f, float inv_norm, uint16_t *__restrict__ act) {
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint16_t a = f2bf(f[v] * inv_norm), b = f2bf(-(f[v] * inv_norm));
    for (int c = 0; c < CH; c++) {
        act[v * CH + c] = c == 0 ? a : (uint16_t)0;
        act[(n + v) * CH + c] = c == 0 ? b : (uint16_t)0;
    }
}

// the correction norm (F(q) - F(-q)) / 2 from the raw outputs F of both halves, clamped to
// +-hs[v] when hs is given
__device__ __forceinline__ float correction(long n, long v, const float *raw, float norm, const float *hs,
                                            bool &clamped) {
    const float c = norm * (0.5f * (raw[v] - raw[n + v]));
    clamped = false;
    if (hs) {
        const float h = hs[v];
        if (c > h || c < -h) {
            clamped = true;
            return c > h ? h : -h;
        }
    }
    return c;
}

// out = q + correction
__global__ void k_combine(long n, const float *__restrict__ q, const float *__restrict__ raw, float norm,
                          const float *hs, float *out) {  // hs may be out
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    bool cl;
    out[v] = q[v] + correction(n, v, raw, norm, hs, cl);
}

// MSE loss gradient in the raw outputs of both halves, +-norm/2 x 2 (out - target) / n, into
// channel 0 of 16-channel bfloat16 (volumes 0..n-1 and n..2n-1), and the per-workgroup partial
// sums of (out - target)^2
__global__ void __launch_bounds__(256) k_mse(long n, const float *__restrict__ out, const float *__restrict__ target,
                                             const float *__restrict__ raw, float norm, const float *__restrict__ hs,
                                             uint16_t *__restrict__ g, float *__restrict__ part) {
    __shared__ float red[256];
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    float e2 = 0.0f;
    if (v < n) {
        const float e = out[v] - target[v];
        e2 = e * e;
        bool cl;
        (void)correction(n, v, raw, norm, hs, cl);
        const float d = cl ? 0.0f : 0.5f * norm * (2.0f * e / (float)n);  // no gradient through the clamp
        for (int c = 0; c < CH; c++) {
            g[v * CH + c] = c == 0 ? f2bf(d) : (uint16_t)0;
            g[(n + v) * CH + c] = c == 0 ? f2bf(-d) : (uint16_t)0;
        }
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// weights of one layer (float32 W[t][co][c]) -> matrix-core B operands: forward Wb[t][co][c],
// transposed Wt[t'][c][co] = W[26 - t'][co][c] (the data gradient's conv)
__global__ void k_prep_weights(const float *__restrict__ w, uint16_t *__restrict__ wb, uint16_t *__restrict__ wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= LAYER_W) return;
    const int t = i / (CH * CH), co = (i / CH) % CH, c = i % CH;
    const uint16_t b = f2bf(w[i]);
    wb[i] = b;
    wt[((NT - 1 - t) * CH + c) * CH + co] = b;
}

__global__ void k_apply_mask(long n, float *__restrict__ g, const float *__restrict__ mask) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[i] *= mask[i];
}

// Adam on the packed parameters; entries with mask 0 (the zero padding) stay 0
__global__ void k_adam(long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, const float *__restrict__ mask, float lr, float b1, float b2, float c1,
                       float c2, float eps) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * mask[i];
    m[i] = b1 * m[i] + (1.0f - b1) * gi;
    v[i] = b2 * v[i] + (1.0f - b2) * gi * gi;
    p[i] = (p[i] - lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps)) * mask[i];
}

// patches: dst[i] = scale x the src cube of P^3 at origins (o1, o2, o3)[i] (fast .. slow) of an
// n1 x n2 x n3 field (z fastest); output volume i is X = Y = Z = P with z fastest
__global__ void k_gather(const float *__restrict__ src, long n1, long n2, long n3, const long *__restrict__ org,
                         long P, long npatch, float scale, float *__restrict__ dst) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = P * P * P;
    if (i >= per * npatch) return;
    const long pi = i / per, e = i - pi * per;
    const long z = e % P, y = (e / P) % P, x = e / (P * P);
    const long o1 = org[3 * pi], o2 = org[3 * pi + 1], o3 = org[3 * pi + 2];
    dst[i] = src[((o3 + x) * n2 + (o2 + y)) * n1 + (o1 + z)] * scale;
}

// ---- the network --------------------------------------------------------------------------
struct Net {
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr, *mask = nullptr;  // NPARAM each
    uint16_t *wb = nullptr, *wt = nullptr;  // 3 layers of B operands each
    long step = 0;
    long cap = 0;  // voxels the workspace holds
    uint16_t *x0 = nullptr, *h1 = nullptr, *h2 = nullptr, *d = nullptr, *d2 = nullptr;
    float *part = nullptr;
    long part_cap = 0;
    float *lossp = nullptr;
    float *raw = nullptr;  // conv3's raw output of both halves
    ~Net() {
        for (void *q : {(void *)p, (void *)g, (void *)m, (void *)v, (void *)mask, (void *)wb, (void *)wt, (void *)x0,
                        (void *)h1, (void *)h2, (void *)d, (void *)d2, (void *)part, (void *)lossp, (void *)raw})
            if (q) (void)hipFree(q);
    }
};

float *layer_w(float *p, int l) { return p + (long)l * LAYER_P; }
float *layer_b(float *p, int l) { return p + (long)l * LAYER_P + LAYER_W; }

// workspace for nvox voxels of input, i.e. 2 nvox of activations (the field and its negation)
long ensure_ws(Net *n, long nvox) {
    if (nvox <= n->cap) return 0;
    for (uint16_t **q : {&n->x0, &n->h1, &n->h2, &n->d, &n->d2}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
        LC_TRY(hipMalloc(q, sizeof(uint16_t) * CH * 2 * nvox));
    }
    if (n->raw) (void)hipFree(n->raw);
    n->raw = nullptr;
    LC_TRY(hipMalloc(&n->raw, sizeof(float) * 2 * nvox));
    n->cap = nvox;
    return 0;
}

long ensure_part(Net *n, long len) {
    if (len <= n->part_cap) return 0;
    if (n->part) (void)hipFree(n->part);
    n->part = nullptr;
    LC_TRY(hipMalloc(&n->part, sizeof(float) * len));
    n->part_cap = len;
    return 0;
}

long prep(Net *n, hipStream_t s) {
    for (int l = 0; l < 3; l++) {
        k_prep_weights<<<(LAYER_W + 255) / 256, 256, 0, s>>>(layer_w(n->p, l), n->wb + (long)l * LAYER_W,
                                                            n->wt + (long)l * LAYER_W);
    }
    LC_TRY(hipGetLastError());
    return 0;
}

unsigned grid_conv(const Vol &v) { return (unsigned)conv_blocks(v); }

// both halves: the volumes and their negations
Vol twice(const Vol &v) { return Vol{2 * v.npatch, v.X, v.Y, v.Z}; }

// out = q + norm f(q / norm) for the volumes q (float32); keeps x0, h1, h2 (of both halves) for
// the backward pass
long forward(Net *n, const Vol &vol, const float *q, float *out, float norm, const float *hs, hipStream_t s) {
    const long nv = vol.nvox();
    const Vol v2 = twice(vol);
    if (ensure_ws(n, nv)) return -1;
    if (prep(n, s)) return -1;
    k_pack1<<<(unsigned)((nv + 255) / 256), 256, 0, s>>>(nv, q, 1.0f / norm, n->x0);
    k_conv<OUT_RELU><<<grid_conv(v2), 256, 0, s>>>(v2, n->x0, n->wb, layer_b(n->p, 0), nullptr, n->h1, nullptr);
    k_conv<OUT_RELU><<<grid_conv(v2), 256, 0, s>>>(v2, n->h1, n->wb + LAYER_W, layer_b(n->p, 1), nullptr, n->h2,
                                                   nullptr);
    k_conv<OUT_RAW><<<grid_conv(v2), 256, 0, s>>>(v2, n->h2, n->wb + 2 * LAYER_W, layer_b(n->p, 2), nullptr, nullptr,
                                                  n->raw);
    k_combine<<<(unsigned)((nv + 255) / 256), 256, 0, s>>>(nv, q, n->raw, norm, hs, out);
    LC_TRY(hipGetLastError());
    return 0;
}

// weight and bias gradients of one layer from its output gradient dO and its input
long layer_grads(Net *n, const Vol &vol, const uint16_t *dO, const uint16_t *in, int l, hipStream_t s) {
    const long nt = wgrad_tiles(vol), nv = vol.nvox();
    const long tpb = 256;  // tiles per workgroup
    const long nblk = (nt + tpb - 1) / tpb;
    const long vpb = 4096;
    const long nbb = (nv + vpb - 1) / vpb;
    if (ensure_part(n, std::max(nblk * 4 * NT * CH * CH, nbb * CH))) return -1;
    k_wgrad<<<dim3((unsigned)nblk, NT), 256, 0, s>>>(vol, dO, in, tpb, n->part);
    k_reduce_parts<<<(LAYER_W + 255) / 256, 256, 0, s>>>(n->part, nblk * 4, LAYER_W, layer_w(n->g, l));
    k_bgrad<<<(unsigned)nbb, 256, 0, s>>>(nv, dO, vpb, n->part);
    k_reduce_parts<<<1, CH, 0, s>>>(n->part, nbb, CH, layer_b(n->g, l));
    LC_TRY(hipGetLastError());
    return 0;
}

long loss_grad(Net *n, const Vol &vol, const float *q, const float *target, float *out, float norm, const float *hs,
               float *loss, hipStream_t s) {
    const long nv = vol.nvox();
    if (forward(n, vol, q, out, norm, hs, s)) return -1;
    const long nb = (nv + 255) / 256;
    if (ensure_part(n, nb)) return -1;
    k_mse<<<(unsigned)nb, 256, 0, s>>>(nv, out, target, n->raw, norm, hs, n->d, n->part);
    k_reduce_parts<<<1, 1, 0, s>>>(n->part, nb, 1, n->lossp);
    // layer 3: dO = d (channel 0), over both halves
    const Vol v2 = twice(vol);
    if (layer_grads(n, v2, n->d, n->h2, 2, s)) return -1;
    k_conv<OUT_MASK><<<grid_conv(v2), 256, 0, s>>>(v2, n->d, n->wt + 2 * LAYER_W, nullptr, n->h2, n->d2, nullptr);
    if (layer_grads(n, v2, n->d2, n->h1, 1, s)) return -1;
    k_conv<OUT_MASK><<<grid_conv(v2), 256, 0, s>>>(v2, n->d2, n->wt + LAYER_W, nullptr, n->h1, n->d, nullptr);
    if (layer_grads(n, v2, n->d, n->x0, 0, s)) return -1;
    k_apply_mask<<<(NPARAM + 255) / 256, 256, 0, s>>>(NPARAM, n->g, n->mask);  // no padding, no biases
    LC_TRY(hipGetLastError());
    if (loss) {
        float l = 0.0f;
        LC_TRY(hipMemcpyAsync(&l, n->lossp, sizeof(float), hipMemcpyDeviceToHost, s));
        LC_TRY(hipStreamSynchronize(s));
        *loss = l / (float)nv;
    }
    return 0;
}

void host_mask(std::vector<float> &mk) {
    mk.assign(NPARAM, 0.0f);
    for (int l = 0; l < 3; l++)
        for (int t = 0; t < NT; t++)
            for (int co = 0; co < CH; co++)
                for (int c = 0; c < CH; c++) {
                    const bool ok = (l == 0 ? c == 0 : true) && (l == 2 ? co == 0 : true);
                    mk[(long)l * LAYER_P + (t * CH + co) * CH + c] = ok ? 1.0f : 0.0f;
                }
    // no biases: f(0) = 0 and f(a q) = a f(q) for a > 0
}

// out = ca a + cb b + cc c, b and c optional, evaluated in that order: the combinations of a
// closed-loop temporal predictor (prediction 2 a - b, residual x - prediction, reconstruction
// prediction + decoded residual), the same kernel on the encoding and the decoding side
// (16-byte accesses over the first n & ~3 elements when every array is 16-byte aligned, then
// the last n % 4 one by one; otherwise element by element -- the same arithmetic per element,
// so the two forms give the same bits)
__device__ __forceinline__ float lin3(float ca, float a, float cb, const float *b, long i, float cc, const float *c) {
    float v = ca * a;
    if (b) v = v + cb * b[i];
    if (c) v = v + cc * c[i];
    return v;
}

__global__ void k_lincomb_scalar(long n, float ca, const float *__restrict__ a, float cb,
                                 const float *__restrict__ b, float cc, const float *__restrict__ c,
                                 float *__restrict__ out) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = lin3(ca, a[i], cb, b, i, cc, c);
}

__global__ void k_lincomb(long n, float ca, const float *__restrict__ a, float cb, const float *__restrict__ b,
                          float cc, const float *__restrict__ c, float *__restrict__ out) {
    const long stride = (long)gridDim.x * blockDim.x;
    const long n4 = n >> 2;
    for (long i4 = (long)blockIdx.x * blockDim.x + threadIdx.x; i4 < n4; i4 += stride) {
        const float4 av = reinterpret_cast<const float4 *>(a)[i4];
        float4 bv = make_float4(0.0f, 0.0f, 0.0f, 0.0f), cv = bv;
        if (b) bv = reinterpret_cast<const float4 *>(b)[i4];
        if (c) cv = reinterpret_cast<const float4 *>(c)[i4];
        float4 o;
        o.x = ca * av.x;
        o.y = ca * av.y;
        o.z = ca * av.z;
        o.w = ca * av.w;
        if (b) {
            o.x = o.x + cb * bv.x;
            o.y = o.y + cb * bv.y;
            o.z = o.z + cb * bv.z;
            o.w = o.w + cb * bv.w;
        }
        if (c) {
            o.x = o.x + cc * cv.x;
            o.y = o.y + cc * cv.y;
            o.z = o.z + cc * cv.z;
            o.w = o.w + cc * cv.w;
        }
        reinterpret_cast<float4 *>(out)[i4] = o;
    }
    for (long i = n4 * 4 + (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = lin3(ca, a[i], cb, b, i, cc, c);
}

// the code's header (magic, grid, block size, levels, digits per word; word 7, the largest
// block scale, is written by k_norm after the blocks), written on the stream so that encoding
// needs no host memory and no wait
__global__ void k_header(uint32_t *code, uint32_t n1, uint32_t n2, uint32_t n3, uint32_t lv) {
    const uint32_t h[8] = {MAGIC, n1, n2, n3, (uint32_t)BS, lv, (uint32_t)q_digits((int)lv), 0u};
    if (threadIdx.x < 8) code[threadIdx.x] = h[threadIdx.x];
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

const char *wavefd_lcodec_LastError(void) { return g_err.c_str(); }

long wavefd_lcodec_combine(long n, float ca, const float *a, float cb, const float *b, float cc, const float *c,
                           float *out, void *stream) {
    if (n < 0 || (n > 0 && (!a || !out))) {
        g_err = "combine: bad arguments";
        return -1;
    }
    if (n == 0) return 0;
    // segment slots of an odd-sized frame sit at 4-byte offsets: 16-byte accesses only when
    // every array allows them
    auto al16 = [](const void *q) { return ((uintptr_t)q & 15u) == 0; };
    const bool vec = al16(a) && al16(out) && (!b || al16(b)) && (!c || al16(c));
    if (vec) {
        const long blocks = std::max<long>(1, std::min<long>((n / 4 + 255) / 256, 1L << 20));
        k_lincomb<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(n, ca, a, cb, b, cc, c, out);
    } else {
        const long blocks = std::max<long>(1, std::min<long>((n + 255) / 256, 1L << 20));
        k_lincomb_scalar<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(n, ca, a, cb, b, cc, c, out);
    }
    LC_TRY(hipGetLastError());
    return 0;
}

long wavefd_lcodec_code_words(long n1, long n2, long n3) {
    if (n1 <= 0 || n2 <= 0 || n3 <= 0) return -1;
    return code_words(qgeo(n1, n2, n3));
}

// levels: 11 (the default), 5 or 3 quantisation levels
long wavefd_lcodec_code_words_l(long n1, long n2, long n3, long levels) {
    if (n1 <= 0 || n2 <= 0 || n3 <= 0 || !q_levels_ok((levels - 1) / 2) || levels % 2 == 0) return -1;
    return code_words(qgeo(n1, n2, n3, (int)((levels - 1) / 2)));
}

long wavefd_lcodec_encode_l(const float *x, long n1, long n2, long n3, long levels, uint32_t *code, void *stream) {
    if (n1 <= 0 || n2 <= 0 || n3 <= 0 || !x || !code || levels % 2 == 0 || !q_levels_ok((levels - 1) / 2)) {
        g_err = "encode: bad arguments (levels 11, 5 or 3)";
        return -1;
    }
    const QGeo g = qgeo(n1, n2, n3, (int)((levels - 1) / 2));
    hipStream_t s = (hipStream_t)stream;
    k_header<<<1, 64, 0, s>>>(code, (uint32_t)n1, (uint32_t)n2, (uint32_t)n3, (uint32_t)g.lv);
    LC_TRY(hipGetLastError());
    LC_TRY(hipMemsetAsync(code + 8, 0, sizeof(uint32_t) * (pay_off(g) - 8), s));  // scales and flags
    const unsigned nwg = (unsigned)(((g.m1 + 3) / 4) * g.m2 * g.m3);
    if (g.lv == 1)
        k_encode<1><<<nwg, 256, 0, s>>>(g, x, code);
    else if (g.lv == 2)
        k_encode<2><<<nwg, 256, 0, s>>>(g, x, code);
    else
        k_encode<5><<<nwg, 256, 0, s>>>(g, x, code);
    LC_TRY(hipGetLastError());
    k_norm<<<1, 1024, 0, s>>>(g, code);
    LC_TRY(hipGetLastError());
    return 0;
}

long wavefd_lcodec_encode(const float *x, long n1, long n2, long n3, uint32_t *code, void *stream) {
    return wavefd_lcodec_encode_l(x, n1, n2, n3, 2 * LV + 1, code, stream);
}

// the grid dimensions stored in a code (device memory)
static long read_header(const uint32_t *code, long *n, hipStream_t s, float *norm = nullptr, int *lv = nullptr) {
    uint32_t hdr[8];
    LC_TRY(hipMemcpyAsync(hdr, code, sizeof(hdr), hipMemcpyDeviceToHost, s));
    LC_TRY(hipStreamSynchronize(s));
    if (hdr[0] != MAGIC || hdr[4] != (uint32_t)BS || !q_levels_ok((long)hdr[5]) ||
        hdr[6] != (uint32_t)q_digits((int)hdr[5])) {
        g_err = "not a learned-codec code";
        return -1;
    }
    n[0] = hdr[1];
    n[1] = hdr[2];
    n[2] = hdr[3];
    if (lv) *lv = (int)hdr[5];
    if (norm) {
        float g;
        memcpy(&g, &hdr[7], 4);
        *norm = g > 0.0f ? g : 1.0f;
    }
    return 0;
}

// the code's largest block scale (the decoder's normalisation; 1 for an all-zero field)
float wavefd_lcodec_norm(const uint32_t *code) {
    long n[3];
    float g = 1.0f;
    if (read_header(code, n, nullptr, &g)) return -1.0f;
    return g;
}

long wavefd_lcodec_dequantize(const uint32_t *code, float *out, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    long n[3];
    int lv = LV;
    if (read_header(code, n, s, nullptr, &lv)) return -1;
    launch_dequantize(qgeo(n[0], n[1], n[2], lv), code, out, s);
    LC_TRY(hipGetLastError());
    return 0;
}

// the plain decode of a code whose grid the caller knows (no header read, no host wait)
long wavefd_lcodec_dequantize_nl(const uint32_t *code, long n1, long n2, long n3, long levels, float *out,
                                 void *stream) {
    if (n1 <= 0 || n2 <= 0 || n3 <= 0 || !code || !out || levels % 2 == 0 || !q_levels_ok((levels - 1) / 2)) {
        g_err = "dequantize_n: bad arguments (levels 11, 5 or 3)";
        return -1;
    }
    launch_dequantize(qgeo(n1, n2, n3, (int)((levels - 1) / 2)), code, out, (hipStream_t)stream);
    LC_TRY(hipGetLastError());
    return 0;
}

long wavefd_lcodec_dequantize_n(const uint32_t *code, long n1, long n2, long n3, float *out, void *stream) {
    return wavefd_lcodec_dequantize_nl(code, n1, n2, n3, 2 * LV + 1, out, stream);
}

// pred = ca a + cb b + cc c and res = x - pred over n floats (device; b, c may be NULL)
long wavefd_lcodec_predict_residual(long n, float ca, const float *a, float cb, const float *b, float cc,
                                    const float *c, const float *x, float *pred, float *res, void *stream) {
    if (n < 0 || (n > 0 && (!a || !x || !pred || !res))) {
        g_err = "predict_residual: bad arguments";
        return -1;
    }
    if (n == 0) return 0;
    const long blocks = std::min<long>((n + 255) / 256, 1L << 20);
    k_pred_res<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(n, ca, a, cb, b, cc, c, x, pred, res);
    LC_TRY(hipGetLastError());
    return 0;
}

// half a quantisation step for every value of the field, s / (levels - 1) (the bound of |x - x_hat|)
long wavefd_lcodec_halfstep(const uint32_t *code, float *out, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    long n[3];
    int lv = LV;
    if (read_header(code, n, s, nullptr, &lv)) return -1;
    const QGeo g = qgeo(n[0], n[1], n[2], lv);
    k_halfstep<<<(unsigned)g.nblocks(), 256, 0, s>>>(g, code, out);
    LC_TRY(hipGetLastError());
    return 0;
}

long wavefd_lcodec_nparams(void) { return NPARAM; }

void *wavefd_lcodec_net_create(void) {
    Net *n = new Net();
    bool ok = true;
    for (float **q : {&n->p, &n->g, &n->m, &n->v, &n->mask}) ok = ok && hipMalloc(q, sizeof(float) * NPARAM) == hipSuccess;
    ok = ok && hipMalloc(&n->wb, sizeof(uint16_t) * 3 * LAYER_W) == hipSuccess;
    ok = ok && hipMalloc(&n->wt, sizeof(uint16_t) * 3 * LAYER_W) == hipSuccess;
    ok = ok && hipMalloc(&n->lossp, sizeof(float)) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        g_err = "net_create: out of device memory";
        delete n;
        return nullptr;
    }
    std::vector<float> mk;
    host_mask(mk);
    std::vector<float> z(NPARAM, 0.0f);
    ok = hipMemcpy(n->mask, mk.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) == hipSuccess;
    for (float *q : {n->p, n->g, n->m, n->v}) ok = ok && hipMemcpy(q, z.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        g_err = "net_create: copy failed";
        delete n;
        return nullptr;
    }
    return n;
}

void wavefd_lcodec_net_free(void *net) { delete reinterpret_cast<Net *>(net); }

// parameters (host arrays of nparams floats): per layer W[t][co][c] (27 x 16 x 16) then the 16
// biases; entries outside the real layer shapes (conv1: c > 0; conv3: co > 0) are ignored/0
long wavefd_lcodec_net_get(void *net, float *params) {
    Net *n = reinterpret_cast<Net *>(net);
    LC_TRY(hipMemcpy(params, n->p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return 0;
}

long wavefd_lcodec_net_set(void *net, const float *params) {
    Net *n = reinterpret_cast<Net *>(net);
    std::vector<float> mk, q(params, params + NPARAM);
    host_mask(mk);
    for (long i = 0; i < NPARAM; i++) q[i] *= mk[i];
    LC_TRY(hipMemcpy(n->p, q.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    std::vector<float> z(NPARAM, 0.0f);
    LC_TRY(hipMemcpy(n->m, z.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    LC_TRY(hipMemcpy(n->v, z.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    n->step = 0;
    return 0;
}

// He-normal initial weights (fan-in 27 x real input channels); the last layer starts at zero so
// that the untrained decoder is the plain dequantisation
long wavefd_lcodec_net_init(void *net, unsigned long seed) {
    std::mt19937_64 rng(seed);
    std::vector<float> q(NPARAM, 0.0f);
    for (int l = 0; l < 2; l++) {
        const float sd = std::sqrt(2.0f / (float)(NT * (l == 0 ? 1 : CH)));
        std::normal_distribution<float> nd(0.0f, sd);
        for (int i = 0; i < LAYER_W; i++) q[(long)l * LAYER_P + i] = nd(rng);
    }
    return wavefd_lcodec_net_set(net, q.data());
}

// out = q + norm f(q / norm) on npatch volumes of X x Y x Z (z fastest), device arrays; the
// correction clamped to +-hs where hs (half steps, one per voxel) is not null
long wavefd_lcodec_forward(void *net, const float *q, long npatch, long X, long Y, long Z, float norm,
                           const float *hs, float *out, void *stream) {
    Net *n = reinterpret_cast<Net *>(net);
    if (npatch <= 0 || X <= 0 || Y <= 0 || Z <= 0 || !(norm > 0.0f)) {
        g_err = "forward: bad volume or norm";
        return -1;
    }
    return forward(n, Vol{npatch, X, Y, Z}, q, out, norm, hs, (hipStream_t)stream);
}

// loss = mean (out - target)^2 and its gradient in the parameters (kept for adam; copied to
// grad, a host array of nparams floats, when grad is not null); out receives the network output
long wavefd_lcodec_grad(void *net, const float *q, const float *target, long npatch, long X, long Y, long Z, float norm,
                        const float *hs, float *out, float *loss, float *grad, void *stream) {
    Net *n = reinterpret_cast<Net *>(net);
    if (npatch <= 0 || X <= 0 || Y <= 0 || Z <= 0 || !(norm > 0.0f)) {
        g_err = "grad: bad volume or norm";
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    if (loss_grad(n, Vol{npatch, X, Y, Z}, q, target, out, norm, hs, loss, s)) return -1;
    if (grad) {
        LC_TRY(hipMemcpyAsync(grad, n->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, s));
        LC_TRY(hipStreamSynchronize(s));
    }
    return 0;
}

long wavefd_lcodec_adam(void *net, float lr, void *stream) {
    Net *n = reinterpret_cast<Net *>(net);
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)n->step), c2 = 1.0f - std::pow(b2, (float)n->step);
    k_adam<<<(NPARAM + 255) / 256, 256, 0, (hipStream_t)stream>>>(NPARAM, n->p, n->g, n->m, n->v, n->mask, lr, b1, b2,
                                                                   c1, c2, 1e-8f);
    LC_TRY(hipGetLastError());
    return 0;
}

// the learned decode of a whole code: dequantise, then out = q_hat + g f(q_hat / g) with g the
// code's largest block scale; scratch: a device array of the grid's size (the dequantised field)
long wavefd_lcodec_decode(void *net, const uint32_t *code, float *scratch, float *out, void *stream) {
    Net *n = reinterpret_cast<Net *>(net);
    hipStream_t s = (hipStream_t)stream;
    long d[3];
    float norm = 1.0f;
    if (read_header(code, d, s, &norm)) return -1;
    if (wavefd_lcodec_dequantize(code, scratch, stream)) return -1;
    // the half steps go to out, which the network reads before it writes its result there
    // (k_combine reads hs[v] and writes out[v] in the same thread)
    if (wavefd_lcodec_halfstep(code, out, stream)) return -1;
    if (forward(n, Vol{1, d[2], d[1], d[0]}, scratch, out, norm, out, s)) return -1;
    const QGeo g = qgeo(d[0], d[1], d[2]);
    k_apply_flags<<<(unsigned)g.nblocks(), 256, 0, s>>>(g, code, scratch, out);
    LC_TRY(hipGetLastError());
    return 0;
}

// encode with the encoder's per-block choice: quantise, run the learned decode, and flag the
// blocks where it is closer to x than the plain dequantisation (so that no block's decode is
// worse than plain); scratch1, scratch2: device arrays of the field's size
long wavefd_lcodec_encode_learned_l(void *net, const float *x, long n1, long n2, long n3, long levels, uint32_t *code,
                                    float *scratch1, float *scratch2, void *stream) {
    Net *n = reinterpret_cast<Net *>(net);
    hipStream_t s = (hipStream_t)stream;
    if (wavefd_lcodec_encode_l(x, n1, n2, n3, levels, code, stream)) return -1;
    float norm = 1.0f;
    long d[3];
    if (read_header(code, d, s, &norm)) return -1;
    if (wavefd_lcodec_dequantize(code, scratch1, stream)) return -1;
    if (wavefd_lcodec_halfstep(code, scratch2, stream)) return -1;
    if (forward(n, Vol{1, n3, n2, n1}, scratch1, scratch2, norm, scratch2, s)) return -1;
    const QGeo g = qgeo(n1, n2, n3);
    k_select<<<(unsigned)g.nblocks(), 256, 0, s>>>(g, x, scratch1, scratch2, code);
    LC_TRY(hipGetLastError());
    return 0;
}

long wavefd_lcodec_encode_learned(void *net, const float *x, long n1, long n2, long n3, uint32_t *code, float *scratch1,
                                  float *scratch2, void *stream) {
    return wavefd_lcodec_encode_learned_l(net, x, n1, n2, n3, 2 * LV + 1, code, scratch1, scratch2, stream);
}

long wavefd_lcodec_gather(const float *src, long n1, long n2, long n3, const long *origins, long npatch, long P,
                          float scale, float *dst, void *stream) {
    if (P <= 0 || npatch <= 0 || P > n1 || P > n2 || P > n3) {
        g_err = "gather: patch larger than the field";
        return -1;
    }
    std::vector<long> org(origins, origins + 3 * npatch);
    for (long i = 0; i < npatch; i++)
        if (org[3 * i] < 0 || org[3 * i] + P > n1 || org[3 * i + 1] < 0 || org[3 * i + 1] + P > n2 ||
            org[3 * i + 2] < 0 || org[3 * i + 2] + P > n3) {
            g_err = "gather: patch outside the field";
            return -1;
        }
    hipStream_t s = (hipStream_t)stream;
    long *dorg = nullptr;
    LC_TRY(hipMalloc((void **)&dorg, sizeof(long) * 3 * npatch));
    hipError_t e = hipMemcpyAsync(dorg, org.data(), sizeof(long) * 3 * npatch, hipMemcpyHostToDevice, s);
    const long tot = P * P * P * npatch;
    if (e == hipSuccess) {
        k_gather<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(src, n1, n2, n3, dorg, P, npatch, scale, dst);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dorg);
    if (e != hipSuccess) {
        g_err = std::string("gather: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

}  // extern "C"
#pragma GCC visibility pop
