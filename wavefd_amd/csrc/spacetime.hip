// libspacetime -- drop-in for the reference's src/spacetime.cc (built by src/CMakeLists.txt:13).
//
// Same entry points, argument order and C types (declared with citations in
// include/spacetime.h).  Each entry point dispatches on where the caller's arrays live:
//   * host arrays (the reference's case: Julia arrays, or the host mirrors the propagator
//     getters return) are processed on the host, with the reference's loop structure and
//     accumulation order, because the data is there (one thread: the result does not depend
//     on the reference's nthreads, which only partitions race-free work);
//   * device arrays (GetDevicePtr, or any hipMalloc'd buffer) are processed by gfx950
//     kernels on the null stream, after all prior work on the device (the propagators'
//     own streams included) and completed before the call returns.  A call with the same
//     points as an earlier one reuses that call's grouped points on the device (a host
//     comparison instead of the sort and the upload), and host data arrays move only the
//     one time sample the call reads or writes.
// Injection / extraction accumulate per key in the reference's partition order: all points
// sharing a grid index (injection) or a trace (extraction) sit in one partition in order
// (src/spacetime.jl:1087-1129), and the device kernels give one thread per key group.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

struct SourcePoint32 {
    long iu;
    long ir;
    float c;
};
struct SourcePoint64 {
    long iu;
    long ir;
    double c;
};

std::string g_err;

void report(const std::string &s) {
    g_err = s;
    std::fprintf(stderr, "[wavefd_amd spacetime] %s\n", s.c_str());
}

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

// data value of trace ir (1-based) at this call's time sample: data[(ir - 1) * stride + col]
// (a device data array: stride nt, col it - 1; the host's one column staged on the device:
// stride 1, col 0)
template <class P, class T>
__global__ void k_inject(T *__restrict__ field, const T *__restrict__ data, long stride, long col,
                         const P *__restrict__ pts, const int *__restrict__ gs, int ng) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ng) return;
    const long iu = pts[gs[gi]].iu - 1;
    T acc = field[iu];
    for (int i = gs[gi]; i < gs[gi + 1]; i++) acc += pts[i].c * data[(pts[i].ir - 1) * stride + col];
    field[iu] = acc;
}

template <class P, class T>
__global__ void k_extract(T *__restrict__ data, const T *__restrict__ field, long stride, long col,
                          const P *__restrict__ pts, const int *__restrict__ gs, int ng) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ng) return;
    const long id = (pts[gs[gi]].ir - 1) * stride + col;
    T acc = data[id];
    for (int i = gs[gi]; i < gs[gi + 1]; i++) acc += pts[i].c * field[pts[i].iu - 1];
    data[id] = acc;
}

// The partitions' points grouped by key (stable, so per-key order is the partition order) on
// one device: the layout the kernels read.  Kept for later calls with the same points (compared
// on the host), up to LAYOUT_CACHE per point type; the oldest is freed first.
template <class P>
struct Layout {
    bool inject;
    int dev;
    std::vector<P> key;  // the points in partition order, as the caller passed them
    long maxiu, maxir;
    int ng;
    P *dpts;
    int *dgs;
    void *dcol;  // maxir data values: a host data array's column, staged (allocated at first use)
};
constexpr size_t LAYOUT_CACHE = 8;
std::mutex g_layout_mu;

template <class P>
std::vector<Layout<P>> &layouts() {
    static std::vector<Layout<P>> c;
    return c;
}

template <class P>
void free_layout(Layout<P> &l) {
    if (l.dpts) (void)hipFree(l.dpts);
    if (l.dgs) (void)hipFree(l.dgs);
    if (l.dcol) (void)hipFree(l.dcol);
    l.dpts = nullptr;
    l.dgs = nullptr;
    l.dcol = nullptr;
}

// the layout of `flat` (caller's order) for this device: cached, or built and uploaded
template <class P>
Layout<P> *layout_of(bool inject, int dev, std::vector<P> &flat, hipError_t &e) {
    e = hipSuccess;
    auto &c = layouts<P>();
    // members compared (not the struct's padding bytes); coefficients bit for bit
    auto same = [](const std::vector<P> &a, const std::vector<P> &b) {
        if (a.size() != b.size()) return false;
        for (size_t i = 0; i < a.size(); i++)
            if (a[i].iu != b[i].iu || a[i].ir != b[i].ir || std::memcmp(&a[i].c, &b[i].c, sizeof(a[i].c)) != 0)
                return false;
        return true;
    };
    for (auto &l : c)
        if (l.inject == inject && l.dev == dev && same(l.key, flat)) return &l;
    Layout<P> l{inject, dev, flat, 0, 0, 0, nullptr, nullptr, nullptr};
    std::stable_sort(flat.begin(), flat.end(),
                     [inject](const P &a, const P &b) { return inject ? a.iu < b.iu : a.ir < b.ir; });
    std::vector<int> gs;
    for (size_t i = 0; i < flat.size(); i++) {
        const long key = inject ? flat[i].iu : flat[i].ir;
        if (i == 0 || key != (inject ? flat[i - 1].iu : flat[i - 1].ir)) gs.push_back((int)i);
        l.maxiu = std::max(l.maxiu, flat[i].iu);
        l.maxir = std::max(l.maxir, flat[i].ir);
    }
    gs.push_back((int)flat.size());
    l.ng = (int)gs.size() - 1;
    e = hipMalloc(&l.dpts, sizeof(P) * flat.size());
    if (e == hipSuccess) e = hipMalloc(&l.dgs, sizeof(int) * gs.size());
    if (e == hipSuccess) e = hipMemcpy(l.dpts, flat.data(), sizeof(P) * flat.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(l.dgs, gs.data(), sizeof(int) * gs.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_layout(l);
        return nullptr;
    }
    if (c.size() >= LAYOUT_CACHE) {
        free_layout(c.front());
        c.erase(c.begin());
    }
    c.push_back(std::move(l));
    return &c.back();
}

// Inject or extract with the field in device memory: after all prior work on the device, on the
// null stream, completed before returning.  Data on the host moves one time sample (the column
// the call reads, and for extraction writes back); data on the device is read in place.
template <class P, class T>
void device_scatter_gather(bool inject, T *field, T *data, size_t it, size_t nt, P **parts, size_t nparts,
                           const long *npts) {
    // the device field's allocation bounds the grid index: reject points that would land outside
    hipDeviceptr_t base = nullptr;
    size_t span = 0;
    if (hipMemGetAddressRange(&base, &span, (hipDeviceptr_t)field) != hipSuccess) {
        (void)hipGetLastError();
        report("cannot determine the size of the device field");
        return;
    }
    const long ncells = (long)((span - (size_t)((char *)field - (char *)base)) / sizeof(T));
    std::vector<P> flat;
    for (size_t i = 0; i < nparts; i++)
        for (long j = 0; j < npts[i]; j++) {
            const P &q = parts[i][j];
            if (q.iu < 1 || q.iu > ncells || q.ir < 1) {
                report("point outside the device field (iu=" + std::to_string(q.iu) + ")");
                return;
            }
            flat.push_back(q);
        }
    if (flat.empty()) return;
    if (it < 1 || it > nt) {
        report(std::string(inject ? "injectdata" : "extractdata") + ": time index " + std::to_string(it) +
               " outside 1.." + std::to_string(nt));
        return;
    }
    const bool data_dev = is_device(data);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(g_layout_mu);
    Layout<P> *l = e == hipSuccess ? layout_of<P>(inject, dev, flat, e) : nullptr;
    if (l && data_dev) {  // the device data array must hold every trace the points name
        hipDeviceptr_t dbase = nullptr;
        size_t dspan = 0;
        if (hipMemGetAddressRange(&dbase, &dspan, (hipDeviceptr_t)data) != hipSuccess) {
            (void)hipGetLastError();
            report("cannot determine the size of the device data array");
            return;
        }
        const size_t avail = (dspan - (size_t)((char *)data - (char *)dbase)) / sizeof(T);
        if ((size_t)(l->maxir - 1) * nt + it > avail) {
            report("trace " + std::to_string(l->maxir) + " outside the device data array");
            return;
        }
    }
    T *dcol = nullptr;
    std::vector<T> col;
    if (l && e == hipSuccess && !data_dev) {  // the host data's one column, staged
        if (!l->dcol) e = hipMalloc(&l->dcol, sizeof(T) * (size_t)l->maxir);
        dcol = static_cast<T *>(l->dcol);
        col.resize((size_t)l->maxir);
        for (long r = 0; r < l->maxir; r++) col[(size_t)r] = data[(size_t)r * nt + it - 1];
        if (e == hipSuccess) e = hipMemcpy(dcol, col.data(), sizeof(T) * col.size(), hipMemcpyHostToDevice);
    }
    // work already queued on the device (a propagator's stream included) comes first
    if (l && e == hipSuccess) e = hipDeviceSynchronize();
    if (l && e == hipSuccess) {
        T *dd = data_dev ? data : dcol;
        const long stride = data_dev ? (long)nt : 1, c0 = data_dev ? (long)it - 1 : 0;
        if (inject)
            k_inject<P, T><<<(l->ng + 255) / 256, 256>>>(field, dd, stride, c0, l->dpts, l->dgs, l->ng);
        else
            k_extract<P, T><<<(l->ng + 255) / 256, 256>>>(dd, field, stride, c0, l->dpts, l->dgs, l->ng);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e == hipSuccess && !inject && !data_dev) {
            e = hipMemcpy(col.data(), dcol, sizeof(T) * col.size(), hipMemcpyDeviceToHost);
            if (e == hipSuccess)
                for (long r = 0; r < l->maxir; r++) data[(size_t)r * nt + it - 1] = col[(size_t)r];
        }
    }
    if (e != hipSuccess) report(std::string(inject ? "injectdata" : "extractdata") + ": " + hipGetErrorString(e));
}

// ---- injection without allocation (injectdata_float_noalloc) ---------------------------
// For callers that cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  The partitions' points are staged in a file-scope device
// array, `scratch`, of NOALLOC_MAX_POINTS points and NOALLOC_MAX_PARTS + 1 partition offsets,
// which each GPU the library runs on holds until the process ends; no hipMalloc, hipFree or
// host heap allocation happens per call.  Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU (from several host threads) race.  Nothing in it needs clearing: a call
// writes every entry it reads.
constexpr long NOALLOC_MAX_POINTS = 1L << 20;  // 24 MiB of SourcePoint32
constexpr long NOALLOC_MAX_PARTS = 4096;

struct NoallocScratch {
    SourcePoint32 pts[NOALLOC_MAX_POINTS];
    long off[NOALLOC_MAX_PARTS + 1];  // partition p holds points [off[p], off[p+1])
};
static __device__ __align__(256) NoallocScratch scratch;

constexpr int NOALLOC_MAX_DEVICES = 64;
NoallocScratch *g_scratch_ptr[NOALLOC_MAX_DEVICES];  // hipGetSymbolAddress of `scratch`, per device

// the device address of this device's `scratch` (looked up once per device, with it current)
NoallocScratch *scratch_of(int dev) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            report(std::string("injectdata_float_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = reinterpret_cast<NoallocScratch *>(ptr);
    }
    return g_scratch_ptr[dev];
}

// one thread per point; a point that starts a run of equal iu inside its partition adds the run's
// contributions in partition order (runs of one key are disjoint when partitions are sorted by
// iu, as src/spacetime.jl:1049-1129 builds them and the host checks before this launch)
__global__ void k_inject_runs(float *__restrict__ field, const float *__restrict__ data, long it, long nt,
                              const NoallocScratch *__restrict__ s, long nparts, long npts) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    long lo = 0, hi = nparts;  // the partition p with off[p] <= i < off[p+1]
    while (hi - lo > 1) {
        const long mid = (lo + hi) / 2;
        if (s->off[mid] <= i) lo = mid;
        else hi = mid;
    }
    const long end = s->off[lo + 1];
    const long iu = s->pts[i].iu;
    if (i > s->off[lo] && s->pts[i - 1].iu == iu) return;
    float acc = field[iu - 1];
    for (long j = i; j < end && s->pts[j].iu == iu; j++) acc += s->pts[j].c * data[(s->pts[j].ir - 1) * nt + it - 1];
    field[iu - 1] = acc;
}

// one thread per partition, its points in order: the reference's loop (src/spacetime.cc:144-162)
// for partitions not sorted by iu
__global__ void k_inject_parts(float *__restrict__ field, const float *__restrict__ data, long it, long nt,
                               const NoallocScratch *__restrict__ s, long nparts) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nparts) return;
    for (long j = s->off[p]; j < s->off[p + 1]; j++) {
        const SourcePoint32 q = s->pts[j];
        field[q.iu - 1] += q.c * data[(q.ir - 1) * nt + it - 1];
    }
}

// number of floats from ptr to the end of its device allocation, or -1
long device_extent(const void *ptr) {
    hipDeviceptr_t base = nullptr;
    size_t span = 0;
    if (hipMemGetAddressRange(&base, &span, (hipDeviceptr_t)ptr) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    return (long)((span - (size_t)((const char *)ptr - (const char *)base)) / sizeof(float));
}

void inject_noalloc(float *field, float *data, size_t it, size_t nt, SourcePoint32 **parts, size_t nparts,
                    const long *npts) {
    const char *who = "injectdata_float_noalloc: ";
    if (nparts > (size_t)NOALLOC_MAX_PARTS) {
        report(std::string(who) + "more than " + std::to_string(NOALLOC_MAX_PARTS) + " partitions (" +
               std::to_string(nparts) + ")");
        return;
    }
    long total = 0;
    for (size_t i = 0; i < nparts; i++) {
        if (npts[i] < 0) {
            report(std::string(who) + "negative point count");
            return;
        }
        total += npts[i];
        if (total > NOALLOC_MAX_POINTS) {
            report(std::string(who) + "more than " + std::to_string(NOALLOC_MAX_POINTS) + " points");
            return;
        }
    }
    if (total == 0) return;
    hipPointerAttribute_t fa, da;
    std::memset(&fa, 0, sizeof(fa));
    std::memset(&da, 0, sizeof(da));
    int cur = -1;
    if (hipPointerGetAttributes(&fa, field) != hipSuccess || hipPointerGetAttributes(&da, data) != hipSuccess ||
        fa.type != hipMemoryTypeDevice || da.type != hipMemoryTypeDevice || hipGetDevice(&cur) != hipSuccess ||
        fa.device != cur || da.device != cur) {
        (void)hipGetLastError();
        report(std::string(who) + "field and data must be device memory of the current device");
        return;
    }
    const long nfield = device_extent(field), ndata = device_extent(data);
    if (nfield < 0 || ndata < 0 || it < 1 || it > nt) {
        report(std::string(who) + "cannot size the device arrays, or it outside 1..nt");
        return;
    }
    // every access in bounds; runs of equal iu inside each partition contiguous (sorted by iu)
    bool sorted = true;
    for (size_t p = 0; p < nparts; p++)
        for (long j = 0; j < npts[p]; j++) {
            const SourcePoint32 &q = parts[p][j];
            if (q.iu < 1 || q.iu > nfield || q.ir < 1 || (q.ir - 1) * (long)nt + (long)it > ndata) {
                report(std::string(who) + "point outside the device field or data (iu=" + std::to_string(q.iu) +
                       ", ir=" + std::to_string(q.ir) + ")");
                return;
            }
            if (j > 0 && q.iu < parts[p][j - 1].iu) sorted = false;
        }
    NoallocScratch *S = scratch_of(cur);
    if (!S) return;
    // stage the points and the partition offsets (synchronous copies on the null stream, ordered
    // after the previous call's kernel, which read this same array)
    long hoff[NOALLOC_MAX_PARTS + 1];  // (on the stack)
    hoff[0] = 0;
    hipError_t e = hipSuccess;
    for (size_t p = 0; p < nparts; p++) {
        if (e == hipSuccess && npts[p] > 0)
            e = hipMemcpy(&S->pts[hoff[p]], parts[p], sizeof(SourcePoint32) * npts[p], hipMemcpyHostToDevice);
        hoff[p + 1] = hoff[p] + npts[p];
    }
    if (e == hipSuccess) e = hipMemcpy(&S->off[0], hoff, sizeof(long) * (nparts + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (sorted)
            k_inject_runs<<<(unsigned)((total + 255) / 256), 256>>>(field, data, (long)it, (long)nt, S, (long)nparts,
                                                                    total);
        else
            k_inject_parts<<<(unsigned)((nparts + 63) / 64), 64>>>(field, data, (long)it, (long)nt, S, (long)nparts);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) report(std::string(who) + hipGetErrorString(e));
}

// ---- host paths: src/spacetime.cc:144-204 ---------------------------------------------

template <class P, class T>
void host_inject(T *field, const T *data, size_t it, size_t nt, P **parts, size_t nparts, const long *npts,
                 size_t nthreads) {
    for (size_t ip = 0; ip < nparts; ip++)
        for (long j = 0; j < npts[ip]; j++) {
            const P p = parts[ip][j];
            field[p.iu - 1] += p.c * data[(p.ir - 1) * nt + it - 1];
        }
}

template <class P, class T>
void host_extract(T *data, const T *field, P **parts, size_t nparts, const long *npts, size_t it, size_t nt,
                  size_t nthreads) {
    for (size_t ip = 0; ip < nparts; ip++)
        for (long j = 0; j < npts[ip]; j++) {
            const P p = parts[ip][j];
            data[(p.ir - 1) * nt + it - 1] += p.c * field[p.iu - 1];
        }
}

// ---- 8-tap polyphase time interpolation: src/spacetime.cc:4-127 -------------------------
// The reference visits coarse rows in a fixed sequence with trimmed tap ranges at both
// ends: (0,[3,7]) (1,[2,7]) (2,[1,7]) (3..nd-5,[0,7]) (nd-4,[0,6]) (nd-3,[0,5]) (nd-2,[0,4]),
// in size_t arithmetic.  The sequence is kept literally -- including rows visited twice
// when nd < 7 and rows whose index wraps, which the jk >= nm test then skips.  For nd < 7
// the reference also reads/writes d past its end (undefined behaviour); here those taps are
// skipped, which leaves every in-range result unchanged.

struct Row {
    size_t i, lo, hi;
};

inline std::vector<Row> interp_rows(size_t nd, size_t nh, size_t nm) {
    std::vector<Row> rows = {{0, 3, 7}, {1, 2, 7}, {2, 1, 7}};
    // rows with i*nh >= nm touch nothing, so stopping there changes no result (and avoids
    // the reference's near-endless loop when nd < 4 makes nd-4 wrap)
    for (size_t i = 3; i < nd - 4 && i * nh < nm; i++) rows.push_back({i, 0, 7});
    rows.push_back({nd - 4, 0, 6});
    rows.push_back({nd - 3, 0, 5});
    rows.push_back({nd - 2, 0, 4});
    return rows;
}

template <class T>
void host_interp_forward_1d(T **h, T *d, const T *m, size_t nh, size_t nd, size_t nm) {
    std::memset(d, 0, nd * sizeof(T));
    for (const Row &row : interp_rows(nd, nh, nm)) {
        const size_t j = row.i * nh, im3 = row.i - 3;
        for (size_t k = 0; k < nh; k++) {
            const size_t jk = j + k;
            if (jk >= nm) continue;
            for (size_t l = row.lo; l <= row.hi; l++)
                if (im3 + l < nd) d[im3 + l] += h[k][l] * m[jk];
        }
    }
}

template <class T>
void host_interp_adjoint_1d(T **h, T *m, const T *d, size_t nh, size_t nm, size_t nd) {
    std::memset(m, 0, nm * sizeof(T));
    for (const Row &row : interp_rows(nd, nh, nm)) {
        const size_t j = row.i * nh, im3 = row.i - 3;
        for (size_t k = 0; k < nh; k++) {
            const size_t jk = j + k;
            if (jk >= nm) continue;
            for (size_t l = row.lo; l <= row.hi; l++)
                if (im3 + l < nd) m[jk] += h[k][l] * d[im3 + l];
        }
    }
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

void injectdata_float(float *field, float *data, size_t it, size_t nt, SourcePoint32 **partitions,
                      size_t npartitions, long *npoints_per_partition, size_t nthreads) {
    if (is_device(field))
        device_scatter_gather<SourcePoint32, float>(true, field, data, it, nt, partitions, npartitions,
                                                    npoints_per_partition);
    else
        host_inject(field, data, it, nt, partitions, npartitions, npoints_per_partition, nthreads);
}

void injectdata_float_noalloc(float *field, float *data, size_t it, size_t nt, SourcePoint32 **partitions,
                              size_t npartitions, long *npoints_per_partition, size_t nthreads) {
    (void)nthreads;
    inject_noalloc(field, data, it, nt, partitions, npartitions, npoints_per_partition);
}

void injectdata_double(double *field, double *data, size_t it, size_t nt, SourcePoint64 **partitions,
                       size_t npartitions, long *npoints_per_partition, size_t nthreads) {
    if (is_device(field))
        device_scatter_gather<SourcePoint64, double>(true, field, data, it, nt, partitions, npartitions,
                                                     npoints_per_partition);
    else
        host_inject(field, data, it, nt, partitions, npartitions, npoints_per_partition, nthreads);
}

void extractdata_float(float *data, float *field, SourcePoint32 **partitions, size_t npartitions,
                       long *npoints_per_partition, size_t it, size_t nt, size_t nthreads) {
    if (is_device(field))
        device_scatter_gather<SourcePoint32, float>(false, field, data, it, nt, partitions, npartitions,
                                                    npoints_per_partition);
    else
        host_extract(data, field, partitions, npartitions, npoints_per_partition, it, nt, nthreads);
}

void extractdata_double(double *data, double *field, SourcePoint64 **partitions, size_t npartitions,
                        long *npoints_per_partition, size_t it, size_t nt, size_t nthreads) {
    if (is_device(field))
        device_scatter_gather<SourcePoint64, double>(false, field, data, it, nt, partitions, npartitions,
                                                     npoints_per_partition);
    else
        host_extract(data, field, partitions, npartitions, npoints_per_partition, it, nt, nthreads);
}

void interpadjoint_1d_float(float **h, float *m, float *d, size_t nh, size_t nm, size_t nd) {
    host_interp_adjoint_1d(h, m, d, nh, nm, nd);
}
void interpadjoint_1d_double(double **h, double *m, double *d, size_t nh, size_t nm, size_t nd) {
    host_interp_adjoint_1d(h, m, d, nh, nm, nd);
}
void interpforward_1d_float(float **h, float *d, float *m, size_t nh, size_t nd, size_t nm) {
    host_interp_forward_1d(h, d, m, nh, nd, nm);
}
void interpforward_1d_double(double **h, double *d, double *m, size_t nh, size_t nd, size_t nm) {
    host_interp_forward_1d(h, d, m, nh, nd, nm);
}

void interpadjoint_nd_float(float **h, float *m, float *d, size_t nh, size_t nm, size_t nd, size_t nslow,
                            size_t nthreads) {
    for (size_t i = 0; i < nslow; i++) host_interp_adjoint_1d(h, m + i * nm, d + i * nd, nh, nm, nd);
}
void interpadjoint_nd_double(double **h, double *m, double *d, size_t nh, size_t nm, size_t nd, size_t nslow,
                             size_t nthreads) {
    for (size_t i = 0; i < nslow; i++) host_interp_adjoint_1d(h, m + i * nm, d + i * nd, nh, nm, nd);
}
void interpforward_nd_float(float **h, float *d, float *m, size_t nh, size_t nd, size_t nm, size_t nslow,
                            size_t nthreads) {
    std::memset(d, 0, nd * nslow * sizeof(float));
    for (size_t i = 0; i < nslow; i++) host_interp_forward_1d(h, d + i * nd, m + i * nm, nh, nd, nm);
}
void interpforward_nd_double(double **h, double *d, double *m, size_t nh, size_t nd, size_t nm, size_t nslow,
                             size_t nthreads) {
    std::memset(d, 0, nd * nslow * sizeof(double));
    for (size_t i = 0; i < nslow; i++) host_interp_forward_1d(h, d + i * nd, m + i * nm, nh, nd, nm);
}

const char *spacetime_LastError(void) { return g_err.c_str(); }

}  // extern "C"
#pragma GCC visibility pop
