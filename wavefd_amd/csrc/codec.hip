// libwavefd_codec -- lossy block codec for wavefield snapshots (the reference's Compressor in
// its UInt32 mode, src/compressedio.jl:1-148, delegates to the un-vendored CvxCompress; this
// is the MI355X replacement, exact storage stays available through the raw modes).
//
// Format (little-endian 32-bit words):
//   word 0      magic 0x57464443 ("WFDC")
//   words 1..3  n1 n2 n3    (fast .. slow extents)
//   words 4..6  b1 b2 b3    (block extents; the last block along an axis is cut to the grid)
//   word 7      bits        (8 or 16)
//   then one float32 scale per block (block max |x|), blocks in fast-to-slow order,
//   then one uint32 per block: the word offset of its data from the start of the data area,
//   then the quantised values block by block, each value round(x / scale * L) with
//   L = 2^(bits-1) - 1, the block's own elements only (fast to slow inside the block),
//   packed 4 (bits 8) or 2 (bits 16) per word, each block padded to a whole word.
// Error bound: |x - x'| <= scale_block / (2 L) per value.
//
// Kernels: one workgroup per block; max |x| by a workgroup reduction, then quantise and pack.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::string g_err;

bool is_device(const void *ptr) {
    if (!ptr) return false;
    hipPointerAttribute_t a;
    std::memset(&a, 0, sizeof(a));
    if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

struct Geo {
    long n1, n2, n3, b1, b2, b3, m1, m2, m3;  // m: blocks per axis
    int bits;
    __host__ __device__ long nblocks() const { return m1 * m2 * m3; }
    __host__ __device__ int per() const { return bits == 8 ? 4 : 2; }
    // extents of block blk (the last block along an axis is cut to the grid)
    __host__ __device__ void extents(long blk, long &k1, long &k2, long &k3, long &e1, long &e2, long &e3) const {
        k1 = blk % m1;
        k2 = (blk / m1) % m2;
        k3 = blk / (m1 * m2);
        e1 = k1 == m1 - 1 ? n1 - k1 * b1 : b1;
        e2 = k2 == m2 - 1 ? n2 - k2 * b2 : b2;
        e3 = k3 == m3 - 1 ? n3 - k3 * b3 : b3;
    }
    __host__ __device__ long words_of(long blk) const {
        long k1, k2, k3, e1, e2, e3;
        extents(blk, k1, k2, k3, e1, e2, e3);
        return (e1 * e2 * e3 + per() - 1) / per();
    }
    long data_words() const {
        long w = 0;
        for (long b = 0; b < nblocks(); b++) w += words_of(b);
        return w;
    }
    long total_words() const { return 8 + 2 * nblocks() + data_words(); }
};

constexpr int TPB = 256;

// grid index of element e (fast to slow) of a block with origin (o1,o2,o3) and extents e1,e2
__device__ __forceinline__ long elem_index(const Geo &g, long o1, long o2, long o3, long e1, long e2, long e) {
    const long i1 = e % e1, i2 = (e / e1) % e2, i3 = e / (e1 * e2);
    return ((o3 + i3) * g.n2 + (o2 + i2)) * g.n1 + (o1 + i1);
}

__global__ void __launch_bounds__(TPB) k_compress(Geo g, const float *__restrict__ x, uint32_t *__restrict__ out) {
    const long blk = blockIdx.x;
    long k1, k2, k3, e1, e2, e3;
    g.extents(blk, k1, k2, k3, e1, e2, e3);
    const long o1 = k1 * g.b1, o2 = k2 * g.b2, o3 = k3 * g.b3, cnt = e1 * e2 * e3;
    __shared__ float red[TPB];
    float m = 0.0f;
    for (long e = threadIdx.x; e < cnt; e += TPB) m = fmaxf(m, fabsf(x[elem_index(g, o1, o2, o3, e1, e2, e)]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const float scale = red[0];
    float *scales = reinterpret_cast<float *>(out + 8);
    if (threadIdx.x == 0) scales[blk] = scale;
    const int L = (1 << (g.bits - 1)) - 1;
    const float inv = scale > 0.0f ? (float)L / scale : 0.0f;
    const uint32_t *offs = out + 8 + g.nblocks();
    uint32_t *data = out + 8 + 2 * g.nblocks() + offs[blk];
    const int per = g.per();
    const long nw = (cnt + per - 1) / per;
    for (long wi = threadIdx.x; wi < nw; wi += TPB) {
        uint32_t word = 0;
        for (int j = 0; j < per; j++) {
            const long e = wi * per + j;
            int q = 0;
            if (e < cnt) q = (int)rintf(x[elem_index(g, o1, o2, o3, e1, e2, e)] * inv);
            const uint32_t u = g.bits == 8 ? (uint32_t)(q & 0xff) : (uint32_t)(q & 0xffff);
            word |= u << (j * g.bits);
        }
        data[wi] = word;
    }
}

__global__ void __launch_bounds__(TPB) k_decompress(Geo g, const uint32_t *__restrict__ in, float *__restrict__ x) {
    const long blk = blockIdx.x;
    long k1, k2, k3, e1, e2, e3;
    g.extents(blk, k1, k2, k3, e1, e2, e3);
    const long o1 = k1 * g.b1, o2 = k2 * g.b2, o3 = k3 * g.b3, cnt = e1 * e2 * e3;
    const float scale = reinterpret_cast<const float *>(in + 8)[blk];
    const int L = (1 << (g.bits - 1)) - 1;
    const uint32_t *data = in + 8 + 2 * g.nblocks() + in[8 + g.nblocks() + blk];
    const int per = g.per();
    for (long e = threadIdx.x; e < cnt; e += TPB) {
        const long idx = elem_index(g, o1, o2, o3, e1, e2, e);
        const uint32_t word = data[e / per];
        const int sh = (int)(e % per) * g.bits;
        int q;
        if (g.bits == 8)
            q = (int)(int8_t)((word >> sh) & 0xff);
        else
            q = (int)(int16_t)((word >> sh) & 0xffff);
        x[idx] = (float)q * scale / (float)L;
    }
}

bool make_geo(const long *n, const long *b, int bits, Geo &g) {
    if (bits != 8 && bits != 16) return false;
    for (int i = 0; i < 3; i++)
        if (n[i] <= 0 || b[i] <= 0) return false;
    g.n1 = n[0]; g.n2 = n[1]; g.n3 = n[2];
    g.b1 = b[0]; g.b2 = b[1]; g.b3 = b[2];
    g.m1 = (g.n1 + g.b1 - 1) / g.b1;
    g.m2 = (g.n2 + g.b2 - 1) / g.b2;
    g.m3 = (g.n3 + g.b3 - 1) / g.b3;
    g.bits = bits;
    return true;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

// words needed for a compressed volume (an upper bound for the output buffer)
long wavefd_codec_words(const long *n, const long *b, long bits) {
    Geo g;
    if (!make_geo(n, b, (int)bits, g)) return -1;
    return g.total_words();
}

// compress field (n1*n2*n3 floats, fast to slow; host or device) into out (host or device,
// >= wavefd_codec_words words); returns the number of words written, -1 on error
long wavefd_codec_compress(const float *field, const long *n, const long *b, long bits, uint32_t *out) {
    Geo g;
    if (!make_geo(n, b, (int)bits, g)) {
        g_err = "bad codec geometry";
        return -1;
    }
    const long nval = g.n1 * g.n2 * g.n3, nw = g.total_words();
    const float *dx = field;
    uint32_t *dout = out;
    float *tx = nullptr;
    uint32_t *tout = nullptr;
    hipError_t e = hipSuccess;
    if (!is_device(field)) {
        e = hipMalloc(&tx, sizeof(float) * nval);
        if (e == hipSuccess) e = hipMemcpy(tx, field, sizeof(float) * nval, hipMemcpyHostToDevice);
        dx = tx;
    }
    const bool out_dev = is_device(out);
    if (e == hipSuccess && !out_dev) {
        e = hipMalloc(&tout, sizeof(uint32_t) * nw);
        dout = tout;
    }
    if (e == hipSuccess) {
        // header, then (after the scales the kernel writes) the block offset table
        std::vector<uint32_t> hdr(8 + 2 * g.nblocks());
        const uint32_t h8[8] = {0x57464443u, (uint32_t)g.n1, (uint32_t)g.n2, (uint32_t)g.n3,
                                (uint32_t)g.b1, (uint32_t)g.b2, (uint32_t)g.b3, (uint32_t)g.bits};
        std::memcpy(hdr.data(), h8, sizeof(h8));
        uint32_t off = 0;
        for (long b = 0; b < g.nblocks(); b++) {
            hdr[8 + g.nblocks() + b] = off;
            off += (uint32_t)g.words_of(b);
        }
        e = hipMemcpy(dout, hdr.data(), sizeof(uint32_t) * hdr.size(), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        k_compress<<<(unsigned)g.nblocks(), TPB>>>(g, dx, dout);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    if (e == hipSuccess && !out_dev) e = hipMemcpy(out, dout, sizeof(uint32_t) * nw, hipMemcpyDeviceToHost);
    if (tx) (void)hipFree(tx);
    if (tout) (void)hipFree(tout);
    if (e != hipSuccess) {
        g_err = std::string("codec compress: ") + hipGetErrorString(e);
        return -1;
    }
    return nw;
}

// decompress words (host or device) into field (host or device); returns 0, -1 on error
long wavefd_codec_decompress(const uint32_t *in, long nwords, float *field) {
    uint32_t hdr[8];
    const bool in_dev = is_device(in);
    if (nwords < 8 || (in_dev ? hipMemcpy(hdr, in, sizeof(hdr), hipMemcpyDeviceToHost) != hipSuccess
                              : (std::memcpy(hdr, in, sizeof(hdr)), false))) {
        g_err = "codec: short input";
        return -1;
    }
    long n[3] = {(long)hdr[1], (long)hdr[2], (long)hdr[3]}, b[3] = {(long)hdr[4], (long)hdr[5], (long)hdr[6]};
    Geo g;
    if (hdr[0] != 0x57464443u || !make_geo(n, b, (int)hdr[7], g) || nwords < g.total_words()) {
        g_err = "codec: bad header or truncated stream";
        return -1;
    }
    const long nval = g.n1 * g.n2 * g.n3, nw = g.total_words();
    // the block offset table is fixed by the geometry: check it before any kernel follows it
    {
        const long nb = g.nblocks();
        std::vector<uint32_t> offs(nb);
        if (in_dev ? hipMemcpy(offs.data(), in + 8 + nb, sizeof(uint32_t) * nb, hipMemcpyDeviceToHost) != hipSuccess
                   : (std::memcpy(offs.data(), in + 8 + nb, sizeof(uint32_t) * nb), false)) {
            g_err = "codec: cannot read the block table";
            return -1;
        }
        uint32_t off = 0;
        for (long b = 0; b < nb; b++) {
            if (offs[b] != off) {
                g_err = "codec: corrupt block table";
                return -1;
            }
            off += (uint32_t)g.words_of(b);
        }
    }
    const uint32_t *din = in;
    uint32_t *tin = nullptr;
    float *tx = nullptr;
    float *dx = field;
    hipError_t e = hipSuccess;
    if (!in_dev) {
        e = hipMalloc(&tin, sizeof(uint32_t) * nw);
        if (e == hipSuccess) e = hipMemcpy(tin, in, sizeof(uint32_t) * nw, hipMemcpyHostToDevice);
        din = tin;
    }
    const bool f_dev = is_device(field);
    if (e == hipSuccess && !f_dev) {
        e = hipMalloc(&tx, sizeof(float) * nval);
        dx = tx;
    }
    if (e == hipSuccess) {
        k_decompress<<<(unsigned)g.nblocks(), TPB>>>(g, din, dx);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    if (e == hipSuccess && !f_dev) e = hipMemcpy(field, dx, sizeof(float) * nval, hipMemcpyDeviceToHost);
    if (tin) (void)hipFree(tin);
    if (tx) (void)hipFree(tx);
    if (e != hipSuccess) {
        g_err = std::string("codec decompress: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

const char *wavefd_codec_LastError(void) { return g_err.c_str(); }

}  // extern "C"
#pragma GCC visibility pop
