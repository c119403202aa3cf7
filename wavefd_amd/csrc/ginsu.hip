// libwavefd_ginsu -- ginsu windows of a large model for per-shot propagation, and the shared
// image they accumulate into.  The reference has no windowing (src/WaveFD.jl:52 leaves it to
// the callers); the rules here are this library's own:
//
//   * every array is a Julia-ordered (nz, ny, nx) float32 field, index k = (ix*ny + iy)*nz + iz
//     (the propagators' layout, src/prop3DAcoIsoDenQ_DEO2_FDTD.h:131);
//   * a window is the box [oz, oz+snz) x [oy, oy+sny) x [ox, ox+snx) of the big grid, wholly
//     inside it; extraction copies it cell for cell (bit-exact);
//   * accumulation adds a sub-box [lz0,lz1) x [ly0,ly1) x [lx0,lx1) of a window-sized array
//     into the big array at the window's offset, one float add per cell, in a fixed order per
//     call: calls made in the same order give the same image bit for bit.
//
// All pointers are device pointers; the work is ordered on the caller's stream (a propagator's
// GetStream, or 0 for the null stream).  Small device helpers for callers without a device
// allocator of their own (the Born driver) live here too.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <string>

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

constexpr int TZ = 64, TY = 4;

// one thread per window cell: z along the lanes (coalesced on both sides), y across the waves
__global__ void __launch_bounds__(TZ *TY) k_extract(const float *__restrict__ big, long bnz, long bny, long oz, long oy,
                                                    long ox, float *__restrict__ sub, long snz, long sny) {
    const long iz = (long)blockIdx.x * TZ + threadIdx.x;
    const long iy = (long)blockIdx.y * TY + threadIdx.y;
    const long ix = blockIdx.z;
    if (iz >= snz || iy >= sny) return;
    sub[(ix * sny + iy) * snz + iz] = big[((ox + ix) * bny + (oy + iy)) * bnz + (oz + iz)];
}

__global__ void __launch_bounds__(TZ *TY) k_accumulate(float *__restrict__ big, long bnz, long bny, long oz, long oy,
                                                       long ox, const float *__restrict__ sub, long snz, long sny,
                                                       long lz0, long lz1, long ly0, long ly1, long lx0) {
    const long iz = lz0 + (long)blockIdx.x * TZ + threadIdx.x;
    const long iy = ly0 + (long)blockIdx.y * TY + threadIdx.y;
    const long ix = lx0 + blockIdx.z;
    if (iz >= lz1 || iy >= ly1) return;
    float &dst = big[((ox + ix) * bny + (oy + iy)) * bnz + (oz + iz)];
    dst = dst + sub[(ix * sny + iy) * snz + iz];
}

bool check_window(long bnz, long bny, long bnx, long oz, long oy, long ox, long snz, long sny, long snx) {
    if (bnz <= 0 || bny <= 0 || bnx <= 0 || snz <= 0 || sny <= 0 || snx <= 0) {
        g_err = "ginsu: non-positive extent";
        return false;
    }
    if (oz < 0 || oy < 0 || ox < 0 || oz + snz > bnz || oy + sny > bny || ox + snx > bnx) {
        g_err = "ginsu: window outside the big grid";
        return false;
    }
    if (snx > 65535L || (snz + TZ - 1) / TZ > 2147483647L || (sny + TY - 1) / TY > 65535L) {
        g_err = "ginsu: window too large for one launch";
        return false;
    }
    return true;
}

long launched(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_err = std::string(what) + ": " + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

// sub = big[window]; returns 0, -1 on error (wavefd_ginsu_LastError)
long wavefd_ginsu_extract(const float *big, long bnz, long bny, long bnx, long oz, long oy, long ox, float *sub,
                          long snz, long sny, long snx, void *stream) {
    if (!check_window(bnz, bny, bnx, oz, oy, ox, snz, sny, snx)) return -1;
    if (!is_device(big) || !is_device(sub)) {
        g_err = "ginsu: extract needs device arrays";
        return -1;
    }
    dim3 grid((unsigned)((snz + TZ - 1) / TZ), (unsigned)((sny + TY - 1) / TY), (unsigned)snx);
    k_extract<<<grid, dim3(TZ, TY), 0, (hipStream_t)stream>>>(big, bnz, bny, oz, oy, ox, sub, snz, sny);
    return launched("ginsu extract");
}

// big[window][l-box] += sub[l-box]; returns 0, -1 on error
long wavefd_ginsu_accumulate(float *big, long bnz, long bny, long bnx, long oz, long oy, long ox, const float *sub,
                             long snz, long sny, long snx, long lz0, long lz1, long ly0, long ly1, long lx0, long lx1,
                             void *stream) {
    if (!check_window(bnz, bny, bnx, oz, oy, ox, snz, sny, snx)) return -1;
    if (lz0 < 0 || ly0 < 0 || lx0 < 0 || lz1 > snz || ly1 > sny || lx1 > snx || lz0 > lz1 || ly0 > ly1 || lx0 > lx1) {
        g_err = "ginsu: accumulation box outside the window";
        return -1;
    }
    if (!is_device(big) || !is_device(sub)) {
        g_err = "ginsu: accumulate needs device arrays";
        return -1;
    }
    if (lz0 == lz1 || ly0 == ly1 || lx0 == lx1) return 0;
    dim3 grid((unsigned)((lz1 - lz0 + TZ - 1) / TZ), (unsigned)((ly1 - ly0 + TY - 1) / TY), (unsigned)(lx1 - lx0));
    k_accumulate<<<grid, dim3(TZ, TY), 0, (hipStream_t)stream>>>(big, bnz, bny, oz, oy, ox, sub, snz, sny, lz0, lz1,
                                                                   ly0, ly1, lx0);
    return launched("ginsu accumulate");
}

// device helpers for drivers that hold no device allocator of their own
void *wavefd_device_alloc(long bytes) {
    void *p = nullptr;
    if (bytes <= 0 || hipMalloc(&p, (size_t)bytes) != hipSuccess) {
        (void)hipGetLastError();
        g_err = "device alloc of " + std::to_string(bytes) + " bytes failed";
        return nullptr;
    }
    return p;
}

void wavefd_device_free(void *p) {
    if (p) (void)hipFree(p);
}

// stream-ordered byte copy between any host/device pointers (kind deduced by the runtime)
long wavefd_device_copy(void *dst, const void *src, long bytes, void *stream) {
    if (bytes < 0) return -1;
    if (bytes == 0) return 0;
    const hipError_t e = hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDefault, (hipStream_t)stream);
    if (e != hipSuccess) {
        g_err = std::string("device copy: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

// stream-ordered zero fill of a device range
long wavefd_device_zero(void *dst, long bytes, void *stream) {
    if (bytes < 0) return -1;
    if (bytes == 0) return 0;
    const hipError_t e = hipMemsetAsync(dst, 0, (size_t)bytes, (hipStream_t)stream);
    if (e != hipSuccess) {
        g_err = std::string("device zero: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

long wavefd_device_sync(void *stream) {
    const hipError_t e = stream ? hipStreamSynchronize((hipStream_t)stream) : hipDeviceSynchronize();
    if (e != hipSuccess) {
        g_err = std::string("sync: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

const char *wavefd_ginsu_LastError(void) { return g_err.c_str(); }

}  // extern "C"
#pragma GCC visibility pop
