// wfd_runtime.hpp -- host runtime behind the WaveFD 3D propagator C-ABI.
//
// One Prop object per handle returned by Prop3DAco*_alloc.  It owns:
//   * the reference's state arrays as device allocations in HBM (layout identical to the
//     reference, src/prop3DAcoIsoDenQ_DEO2_FDTD.h:73-84), zeroed at allocation as the
//     reference's numaFirstTouch does (:94-205);
//   * a pinned host mirror per array, created the first time the array's getter is called;
//   * one HIP stream (default priority) on which every kernel of the handle runs;
//   * lazily allocated device scratch for the two-pass operators and Born paths.
//
// Host coherence (the getters hand out host pointers, src/prop3DAcoIsoDenQ_DEO2_FDTD.jl:58-61,
// which callers read and write between calls, test/spacetime.jl:486-492):
//   mode COHERENT (default): before a call, every exposed array the call reads is copied
//     host->device; after it, every exposed array the call writes is copied device->host and
//     the call returns only when the work is done.  This is the reference's synchronous,
//     host-visible behaviour, at PCIe cost.
//   mode DEVICE: no automatic copies; calls are stream-ordered (asynchronous).  Host mirrors
//     are synchronised only by the explicit SyncToDevice / SyncToHost calls; device pointers
//     come from GetDevicePtr.  This is the fast path.
// Caller-owned arrays passed to Born calls may be host or device memory: host arrays are
// staged through device scratch (and written back when the call accumulates into them),
// device arrays are used in place.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "wfd_born.hpp"
#include "wfd_kernels.hpp"

namespace wfd {

enum FieldId {
    F_V, F_B, F_Q, F_EPS, F_ETA, F_F, F_ST, F_CT, F_SP, F_CP,
    F_PCUR, F_POLD, F_MCUR, F_MOLD, F_PSPACE, F_MSPACE, F_COUNT
};

static const char *const FIELD_NAMES[F_COUNT] = {
    "v", "b", "dtOmegaInvQ", "eps", "eta", "f", "sinTheta", "cosTheta", "sinPhi", "cosPhi",
    "pCur", "pOld", "mCur", "mOld", "pSpace", "mSpace"};

enum Physics { ISO = 0, VTI = 1, TTI = 2 };
enum Coherence { DEVICE = 0, COHERENT = 1 };

typedef unsigned Mask;
constexpr Mask bit(int f) { return 1u << f; }

// process-wide last error of this library
static std::string g_last_error;

static inline void set_error(const std::string &s) {
    g_last_error = s;
    std::fprintf(stderr, "[wavefd_amd] %s\n", s.c_str());
    if (std::getenv("WAVEFD_ABORT_ON_ERROR")) std::abort();
}

#define WFD_TRY(call)                                                                            \
    do {                                                                                         \
        hipError_t _e = (call);                                                                  \
        if (_e != hipSuccess) {                                                                  \
            set_error(std::string(#call) + " failed: " + hipGetErrorString(_e) + " (" + __FILE__ + \
                      ":" + std::to_string(__LINE__) + ")");                                     \
            return false;                                                                        \
        }                                                                                        \
    } while (0)

#define WFD_LAUNCHED()                                                                           \
    do {                                                                                         \
        hipError_t _e = hipGetLastError();                                                       \
        if (_e != hipSuccess) {                                                                  \
            set_error(std::string("kernel launch failed: ") + hipGetErrorString(_e) + " (" +      \
                      __FILE__ + ":" + std::to_string(__LINE__) + ")");                          \
            return false;                                                                        \
        }                                                                                        \
    } while (0)

struct Field {
    float *d = nullptr;
    float *h = nullptr;
    bool used = false;
};

// a registered set of injection or extraction points, grouped by key on the device
struct PointSet {
    SourcePoint32 *pts = nullptr;
    int *gstart = nullptr;
    int ngroups = 0;
    long npts = 0;
    long nr = 0;  // 1 + max trace index - 1, i.e. traces referenced
    bool valid() const { return ngroups > 0; }
};

static inline bool build_pointset(SourcePoint32 **parts, size_t nparts, const long *npoints, bool by_trace,
                                  long ncells, hipStream_t s, PointSet &out) {
    std::vector<SourcePoint32> flat;
    for (size_t i = 0; i < nparts; i++)
        for (long j = 0; j < npoints[i]; j++) {
            const SourcePoint32 &q = parts[i][j];
            // device scatters must stay inside the field: reject what would write outside it
            if (q.iu < 1 || q.iu > ncells || q.ir < 1) {
                set_error("point with iu=" + std::to_string(q.iu) + ", ir=" + std::to_string(q.ir) +
                          " is outside the grid of " + std::to_string(ncells) + " cells");
                return false;
            }
            flat.push_back(q);
        }
    // stable grouping by key keeps the per-key order of the reference's partitions
    std::stable_sort(flat.begin(), flat.end(), [by_trace](const SourcePoint32 &a, const SourcePoint32 &b) {
        return by_trace ? a.ir < b.ir : a.iu < b.iu;
    });
    std::vector<int> gstart;
    long nr = 0;
    for (size_t i = 0; i < flat.size(); i++) {
        const long key = by_trace ? flat[i].ir : flat[i].iu;
        const long prev = i ? (by_trace ? flat[i - 1].ir : flat[i - 1].iu) : -1;
        if (i == 0 || key != prev) gstart.push_back((int)i);
        nr = std::max(nr, flat[i].ir);
    }
    gstart.push_back((int)flat.size());
    if (out.pts) hipFree(out.pts);
    if (out.gstart) hipFree(out.gstart);
    out = PointSet();
    if (flat.empty()) return true;
    WFD_TRY(hipMalloc(&out.pts, sizeof(SourcePoint32) * flat.size()));
    WFD_TRY(hipMalloc(&out.gstart, sizeof(int) * gstart.size()));
    WFD_TRY(hipMemcpyAsync(out.pts, flat.data(), sizeof(SourcePoint32) * flat.size(), hipMemcpyHostToDevice, s));
    WFD_TRY(hipMemcpyAsync(out.gstart, gstart.data(), sizeof(int) * gstart.size(), hipMemcpyHostToDevice, s));
    WFD_TRY(hipStreamSynchronize(s));
    out.ngroups = (int)gstart.size() - 1;
    out.npts = (long)flat.size();
    out.nr = nr;
    return true;
}

static inline void free_pointset(PointSet &p) {
    if (p.pts) hipFree(p.pts);
    if (p.gstart) hipFree(p.gstart);
    p = PointSet();
}

// true when ptr can be dereferenced by kernels in place (device or managed memory)
static inline bool is_device_ptr(const void *ptr) {
    if (!ptr) return false;
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

struct Prop {
    Physics phys;
    bool fs;
    long nthread, nx, ny, nz, nsponge, nbx, nby, nbz;
    float dx, dy, dz, dt;
    Grid g;
    long n;
    hipStream_t stream = nullptr;
    int device = 0;
    int coherence = COHERENT;
    Field f[F_COUNT];
    float *tmp[12] = {};    // operator scratch (two-pass temporaries, Born temporaries)
    float *stage[8] = {};   // staging for host-resident caller arrays
    PointSet src, rcv;      // registered injection / extraction points
    const float *src_data = nullptr;  // device copy of the source data (nr x nt)
    float *rcv_data = nullptr;        // device trace buffer (nr x nt)
    long src_nt = 0, rcv_nt = 0, src_nr = 0, rcv_nr = 0;
    // fused time interpolation (nh > 0): data at the coarse rate (src_nt / rcv_nt coarse
    // samples), filters h (nh x 8) on the device, nm fine samples
    float *src_h = nullptr, *rcv_h = nullptr;
    long src_nh = 0, rcv_nh = 0, src_nm = 0, rcv_nm = 0;
    int src_fields = 1, rcv_fields = 1;
    int fused = 1;  // use the fused single-pass step kernels
    int fused_tile = 0;  // fused kernel tile: 0 automatic, 1 64 (z) x 16 (y), 2 32 x 32
    int fused_lockstep = -1;  // rounds in step: -1 automatic, 0 off, 1 on (where it can run)
    long fused_chunk = 0;     // x-planes per chunk: 0 automatic
    int cu_count = 256;  // compute units of the device (sizes the fused kernel's x-chunks)
    unsigned *fused_sync = nullptr;  // per-XCD round counters of the persistent fused launch
    hipEvent_t ev = nullptr;         // orders this handle's stream against a batch's (lazy)
    bool born_field = false;         // Iso: mCur/mOld hold the coupled Born step's second field
    long *it_dev = nullptr;          // device time index of graph-replayed loops (lazy)
    // Q by sponge distance (nsponge + 1 values, the last the interior's) while dtOmegaInvQ is
    // the SetupDtOmegaInvQ profile: the fused step then reads the table instead of the array.
    // Handing out Q's host mirror or device address (a caller may write it, now or later) ends
    // that for good.
    float *qtab = nullptr;
    bool q_profile = false;
    bool q_exposed = false;
    int dim = 3;                     // 2: a 2D propagator (ny = 1, two-pass 2D kernels)

    Mask model_mask() const {
        Mask m = bit(F_V) | bit(F_B) | bit(F_Q);
        if (phys != ISO) m |= bit(F_EPS) | bit(F_ETA) | bit(F_F);
        if (phys == TTI) m |= bit(F_ST) | bit(F_CT);
        if (phys == TTI && dim == 3) m |= bit(F_SP) | bit(F_CP);  // the 2D TTI has no azimuth
        return m;
    }
    Mask wave_mask() const {
        Mask m = bit(F_PCUR) | bit(F_POLD) | bit(F_PSPACE);
        if (phys != ISO) m |= bit(F_MCUR) | bit(F_MOLD) | bit(F_MSPACE);
        return m;
    }
    Mask all_mask() const { return model_mask() | wave_mask(); }

    bool init() {
        if (hipGetDevice(&device) != hipSuccess) {
            (void)hipGetLastError();
            set_error("no HIP device available");
            return false;
        }
        WFD_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
            cu_count = cus;
        else
            (void)hipGetLastError();
        WFD_TRY(hipMalloc(&fused_sync, 64));
        const Mask m = all_mask();
        for (int i = 0; i < F_COUNT; i++) {
            if (!(m & bit(i))) continue;
            f[i].used = true;
            WFD_TRY(hipMalloc(&f[i].d, sizeof(float) * n));
            WFD_TRY(hipMemsetAsync(f[i].d, 0, sizeof(float) * n, stream));
        }
        WFD_TRY(hipStreamSynchronize(stream));
        return true;
    }

    ~Prop() {
        if (stream) hipStreamSynchronize(stream);
        if (fused_sync) (void)hipFree(fused_sync);
        if (qtab) (void)hipFree(qtab);
        if (it_dev) (void)hipFree(it_dev);
        if (ev) (void)hipEventDestroy(ev);
        for (int i = 0; i < F_COUNT; i++) {
            if (f[i].d) hipFree(f[i].d);
            if (f[i].h) hipHostFree(f[i].h);
        }
        for (auto &t : tmp)
            if (t) hipFree(t);
        for (auto &t : stage)
            if (t) hipFree(t);
        free_pointset(src);
        free_pointset(rcv);
        if (src_data) hipFree((void *)src_data);
        if (rcv_data) hipFree(rcv_data);
        if (src_h) hipFree(src_h);
        if (rcv_h) hipFree(rcv_h);
        if (stream) hipStreamDestroy(stream);
    }

    float *scratch(int i) {
        if (!tmp[i]) {
            if (hipMalloc(&tmp[i], sizeof(float) * n) != hipSuccess) {
                (void)hipGetLastError();
                set_error("hipMalloc of scratch field failed");
                return nullptr;
            }
        }
        return tmp[i];
    }

    // ---- host mirrors ------------------------------------------------------------------
    float *host(int i) {
        if (!f[i].used) return nullptr;
        if (i == F_Q) q_profile = false, q_exposed = true;  // the caller may write Q through the mirror
        if (!f[i].h) {
            if (hipHostMalloc(&f[i].h, sizeof(float) * n, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                set_error(std::string("hipHostMalloc of host mirror failed for ") + FIELD_NAMES[i]);
                return nullptr;
            }
            if (!d2h(bit(i)) || hipStreamSynchronize(stream) != hipSuccess) return nullptr;
        } else if (coherence == COHERENT) {
            // the host view is current after every coherent call
        }
        return f[i].h;
    }

    bool h2d(Mask m) {
        for (int i = 0; i < F_COUNT; i++)
            if ((m & bit(i)) && f[i].h) WFD_TRY(hipMemcpyAsync(f[i].d, f[i].h, sizeof(float) * n, hipMemcpyHostToDevice, stream));
        return true;
    }

    bool d2h(Mask m) {
        for (int i = 0; i < F_COUNT; i++)
            if ((m & bit(i)) && f[i].h) WFD_TRY(hipMemcpyAsync(f[i].h, f[i].d, sizeof(float) * n, hipMemcpyDeviceToHost, stream));
        return true;
    }

    // call bracket: reads/writes are masks of state arrays
    bool begin(Mask reads) { return coherence == COHERENT ? h2d(reads) : true; }

    bool end(Mask writes, bool force_sync = false) {
        if (coherence == COHERENT) {
            if (!d2h(writes)) return false;
            WFD_TRY(hipStreamSynchronize(stream));
        } else if (force_sync) {
            WFD_TRY(hipStreamSynchronize(stream));
        }
        return true;
    }

    // ---- caller-owned arrays -------------------------------------------------------------
    struct Arg {
        float *user;
        float *dev;
        bool staged;
        bool out;
    };

    bool arg(float *user, int slot, bool in, bool out, Arg &a) {
        a.user = user;
        a.out = out;
        if (!user) {
            set_error("null array argument");
            return false;
        }
        if (is_device_ptr(user)) {
            a.dev = user;
            a.staged = false;
            return true;
        }
        if (!stage[slot]) WFD_TRY(hipMalloc(&stage[slot], sizeof(float) * n));
        a.dev = stage[slot];
        a.staged = true;
        if (in) WFD_TRY(hipMemcpyAsync(a.dev, user, sizeof(float) * n, hipMemcpyHostToDevice, stream));
        return true;
    }

    bool finish_args(Arg *args, int nargs) {
        bool any = false;
        for (int i = 0; i < nargs; i++)
            if (args[i].staged) {
                any = true;
                if (args[i].out)
                    WFD_TRY(hipMemcpyAsync(args[i].user, args[i].dev, sizeof(float) * n, hipMemcpyDeviceToHost, stream));
            }
        if (any) WFD_TRY(hipStreamSynchronize(stream));
        return true;
    }

    void swap_p() {
        std::swap(f[F_PCUR], f[F_POLD]);
        if (phys != ISO) std::swap(f[F_MCUR], f[F_MOLD]);
    }
};

// ---- common operations ---------------------------------------------------------------

static inline Prop *make_prop(Physics phys, long fs, long nthread, long nx, long ny, long nz, long nsponge, float dx,
                              float dy, float dz, float dt, long nbx, long nby, long nbz, int dim = 3) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || nsponge <= 0) {
        set_error("alloc: nx, ny, nz and nsponge must be positive");
        return nullptr;
    }
    if (fs > 0 && nz < 8 && nx > 8 && ny > 8) {
        // the free-surface rows kz=0..3 read z-neighbours up to kz+4 = 7
        set_error("alloc: a free surface needs nz >= 8");
        return nullptr;
    }
    Prop *p = new Prop();
    p->phys = phys;
    p->dim = dim;
    p->fs = fs > 0;
    p->nthread = nthread;
    p->nx = nx; p->ny = ny; p->nz = nz;
    p->nsponge = nsponge;
    p->nbx = nbx; p->nby = nby; p->nbz = nbz;
    p->dx = dx; p->dy = dy; p->dz = dz; p->dt = dt;
    p->n = nx * ny * nz;
    p->g.nx = nx; p->g.ny = ny; p->g.nz = nz; p->g.nynz = ny * nz;
    p->g.fs = p->fs ? 1 : 0;
    // 1/d in double, stored to float (src/prop3DAcoIsoDenQ_DEO2_FDTD.h:68-70)
    p->g.idx = (float)(1.0 / dx);
    p->g.idy = (float)(1.0 / dy);
    p->g.idz = (float)(1.0 / dz);
    if (const char *e = std::getenv("WAVEFD_FUSED")) p->fused = std::atoi(e);
    if (const char *e = std::getenv("WAVEFD_FUSED_TILE")) p->fused_tile = std::atoi(e);
    if (!p->init()) {
        delete p;
        return nullptr;
    }
    return p;
}

// The reference rejects freqQ < FLT_EPSILON with a message and exit(EXIT_FAILURE)
// (src/absorb.h:34-39, :98-103).  A library must not end its host process: this prints the same
// message, records it for LastError and returns false; the caller's handle is left unchanged.
static inline bool check_freq_q(float freqQ) {
    if (freqQ >= FLT_EPSILON) return true;
    char msg[1000];
    std::snprintf(msg, sizeof(msg), "Error -- freqQ [%f] is too small!\n", freqQ);
    std::perror(msg);
    set_error(std::string("SetupDtOmegaInvQ: ") + msg);
    return false;
}

// src/absorb.h:85-151
static inline bool setup_q(Prop *p, float freqQ, float qMin, float qInterior) {
    if (!p->begin(bit(F_Q))) return false;
    if ((qMin < FLT_EPSILON) && (qInterior < FLT_EPSILON)) {
        std::printf("Warning -- qMin and qMax unphysical, dtOmegaInvQ set to zero!\n");
        // the reference clears sizeof(float*) bytes, i.e. the first two values (src/absorb.h:107)
        const long nzero = std::min<long>(2, p->n);
        p->q_profile = false;
        WFD_TRY(hipMemsetAsync(p->f[F_Q].d, 0, sizeof(float) * nzero, p->stream));
        return p->end(bit(F_Q), true);
    }
    const long ns = p->nsponge;
    std::vector<float> qprof(ns);
    const float lqmin = std::log(qMin);
    const float lqmax = std::log(qInterior);
    for (long ks = 0; ks < ns; ks++) {
        const float dk = (float)(ks) / (float)(ns - 1);
        const float lq = lqmin + dk * (lqmax - lqmin);
        qprof[ks] = std::exp(lq);
    }
    float *dq = nullptr;
    WFD_TRY(hipMalloc(&dq, sizeof(float) * ns));
    WFD_TRY(hipMemcpyAsync(dq, qprof.data(), sizeof(float) * ns, hipMemcpyHostToDevice, p->stream));
    k_setup_q<<<op_grid(p->g), op_block(), 0, p->stream>>>(p->g, ns, p->dt, freqQ, qInterior, dq, p->f[F_Q].d);
    WFD_LAUNCHED();
    WFD_TRY(hipStreamSynchronize(p->stream));
    hipFree(dq);
    // the same values by sponge distance (k_setup_q's expression, IEEE double on both sides)
    std::vector<float> qt(ns + 1);
    const double PI_REF = 3.1415926535897;
    const double num = (double)(p->dt * 2) * PI_REF * (double)freqQ;
    for (long ks = 0; ks < ns; ks++) qt[ks] = (float)(num / (double)qprof[ks]);
    qt[ns] = (float)(num / (double)qInterior);
    if (!p->qtab) WFD_TRY(hipMalloc(&p->qtab, sizeof(float) * (ns + 1)));
    WFD_TRY(hipMemcpy(p->qtab, qt.data(), sizeof(float) * (ns + 1), hipMemcpyHostToDevice));
    p->q_profile = !p->q_exposed;
    return p->end(bit(F_Q), true);
}

static inline bool scale_spatial(Prop *p) {
    const Mask rd = bit(F_V) | bit(F_B) | bit(F_PSPACE) | (p->phys != ISO ? bit(F_MSPACE) : 0u);
    if (!p->begin(rd)) return false;
    k_scale_space<<<pw_grid(p->n), PW_BLOCK, 0, p->stream>>>(p->n, p->f[F_V].d, p->f[F_B].d, p->f[F_PSPACE].d,
                                                             p->phys != ISO ? p->f[F_MSPACE].d : nullptr);
    WFD_LAUNCHED();
    return p->end(bit(F_PSPACE) | bit(F_MSPACE));
}

// generic D+ / D- on device arrays (src/propagatorStaticFunctions.h:317-738)
static inline bool dplus(Prop *p, const float *in, float *o0, float *o1, float *o2) {
    k_dplus<<<op_grid(p->g), op_block(), 0, p->stream>>>(p->g, in, in, in, o0, o1, o2);
    WFD_LAUNCHED();
    return true;
}

static inline bool dminus(Prop *p, const float *i0, const float *i1, const float *i2, float *o0, float *o1,
                          float *o2) {
    k_dminus<<<op_grid(p->g), op_block(), 0, p->stream>>>(p->g, i0, i1, i2, o0, o1, o2);
    WFD_LAUNCHED();
    return true;
}

static inline bool rplus(Prop *p, const float *in, float *o0, float *o1, float *o2) {
    if (p->phys != TTI) return dplus(p, in, o0, o1, o2);
    k_tti_dplus_rot<<<op_grid(p->g), op_block(), 0, p->stream>>>(p->g, in, in, in, p->f[F_ST].d, p->f[F_CT].d,
                                                                 p->f[F_SP].d, p->f[F_CP].d, o0, o1, o2);
    WFD_LAUNCHED();
    return true;
}

static inline bool rminus(Prop *p, const float *i0, const float *i1, const float *i2, float *o0, float *o1,
                          float *o2) {
    if (p->phys != TTI) return dminus(p, i0, i1, i2, o0, o1, o2);
    k_tti_dminus_rot<<<op_grid(p->g), op_block(), 0, p->stream>>>(p->g, i0, i1, i2, p->f[F_ST].d, p->f[F_CT].d,
                                                                  p->f[F_SP].d, p->f[F_CP].d, o0, o1, o2);
    WFD_LAUNCHED();
    return true;
}

// ---- registered injection / extraction ------------------------------------------------

// the time indices it (1-based) the registered points accept: the samples of the data, or the
// fine steps of a fused interpolation
static inline bool registered_it_ok(const Prop *p, long first, long last) {
    if (p->src.valid()) {
        const long n = p->src_nh > 0 ? p->src_nm : p->src_nt;
        if (first < 1 || last > n) {
            set_error("time index out of range of the registered sources (1.." + std::to_string(n) + ")");
            return false;
        }
    }
    if (p->rcv.valid()) {
        const long n = p->rcv_nh > 0 ? p->rcv_nm : p->rcv_nt;
        if (first < 1 || last > n) {
            set_error("time index out of range of the registered receivers (1.." + std::to_string(n) + ")");
            return false;
        }
    }
    return true;
}

// itp (optional, device): the kernels read the time index there (graph replays; the caller has
// checked its range); otherwise it is checked here
static inline bool inject_registered(Prop *p, long it, const long *itp = nullptr) {
    if (!p->src.valid()) return true;
    if (!itp && !registered_it_ok(p, it, it)) return false;
    const int nb = (p->src.ngroups + 255) / 256;
    auto inject = [&](float *field) {
        if (p->src_nh > 0)
            k_inject_groups_interp<<<nb, 256, 0, p->stream>>>(field, p->src_data, it, p->src_nt, p->src_h, p->src_nh,
                                                              p->src_nm, p->src.pts, p->src.gstart, p->src.ngroups,
                                                              itp);
        else
            k_inject_groups<<<nb, 256, 0, p->stream>>>(field, p->src_data, it, p->src_nt, p->src.pts, p->src.gstart,
                                                       p->src.ngroups, itp);
    };
    if (p->src_fields & 1) {
        inject(p->f[F_PCUR].d);
        WFD_LAUNCHED();
    }
    if ((p->src_fields & 2) && (p->phys != ISO || p->born_field)) {
        inject(p->f[F_MCUR].d);
        WFD_LAUNCHED();
    }
    return true;
}

static inline bool extract_registered(Prop *p, long it, const long *itp = nullptr) {
    if (!p->rcv.valid()) return true;
    if (!itp && !registered_it_ok(p, it, it)) return false;
    const int nb = (p->rcv.ngroups + 255) / 256;
    const int fld = (p->rcv_fields == 2 && (p->phys != ISO || p->born_field)) ? F_MCUR : F_PCUR;
    if (p->rcv_nh > 0)
        k_extract_groups_interp<<<nb, 256, 0, p->stream>>>(p->rcv_data, p->f[fld].d, it, p->rcv_nt, p->rcv_h,
                                                           p->rcv_nh, p->rcv_nm, p->rcv.pts, p->rcv.gstart,
                                                           p->rcv.ngroups, itp);
    else
        k_extract_groups<<<nb, 256, 0, p->stream>>>(p->rcv_data, p->f[fld].d, it, p->rcv_nt, p->rcv.pts,
                                                    p->rcv.gstart, p->rcv.ngroups, itp);
    WFD_LAUNCHED();
    return true;
}

static inline bool set_sources(Prop *p, SourcePoint32 **parts, size_t nparts, const long *npts, const float *data,
                               size_t nt, long fields) {
    if (!build_pointset(parts, nparts, npts, false, p->n, p->stream, p->src)) return false;
    if (p->src_data) hipFree((void *)p->src_data);
    p->src_data = nullptr;
    p->src_nt = (long)nt;
    p->src_nh = p->src_nm = 0;
    p->src_fields = (int)fields;
    p->src_nr = p->src.nr;
    if (!p->src.valid()) return true;
    float *d = nullptr;
    const size_t bytes = sizeof(float) * (size_t)p->src_nr * nt;
    WFD_TRY(hipMalloc(&d, bytes));
    WFD_TRY(hipMemcpyAsync(d, data, bytes, is_device_ptr(data) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           p->stream));
    WFD_TRY(hipStreamSynchronize(p->stream));
    p->src_data = d;
    return true;
}

static inline bool set_receivers(Prop *p, SourcePoint32 **parts, size_t nparts, const long *npts, size_t nt,
                                 long fields) {
    if (!build_pointset(parts, nparts, npts, true, p->n, p->stream, p->rcv)) return false;
    if (p->rcv_data) hipFree(p->rcv_data);
    p->rcv_data = nullptr;
    p->rcv_nt = (long)nt;
    p->rcv_nh = p->rcv_nm = 0;
    p->rcv_fields = (int)fields;
    p->rcv_nr = p->rcv.nr;
    if (!p->rcv.valid()) return true;
    const size_t bytes = sizeof(float) * (size_t)p->rcv_nr * nt;
    WFD_TRY(hipMalloc(&p->rcv_data, bytes));
    WFD_TRY(hipMemsetAsync(p->rcv_data, 0, bytes, p->stream));
    WFD_TRY(hipStreamSynchronize(p->stream));
    return true;
}

// the filters of a fused interpolation: nd coarse samples (at least 7), nh >= 1 fine samples
// per coarse one, nm fine samples; h (nh x 8, host or device) copied to the device
static inline bool set_interp(Prop *p, long nd, const float *h, long nh, long nm, float *&dst_h, long &dst_nh,
                              long &dst_nm) {
    if (nd < 7 || nh < 1 || nm < 1 || !h) {
        set_error("interpolated points: need nd >= 7 coarse samples, nh >= 1, nm >= 1 and filters h");
        return false;
    }
    if (dst_h) hipFree(dst_h);
    dst_h = nullptr;
    WFD_TRY(hipMalloc(&dst_h, sizeof(float) * 8 * nh));
    WFD_TRY(hipMemcpyAsync(dst_h, h, sizeof(float) * 8 * nh, is_device_ptr(h) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           p->stream));
    WFD_TRY(hipStreamSynchronize(p->stream));
    dst_nh = nh;
    dst_nm = nm;
    return true;
}

// sources whose data (nr x nd) is at the coarse rate: each injection at fine step it adds the
// reference's interpadjoint of the data at it (src/spacetime.cc:4-62)
static inline bool set_sources_interp(Prop *p, SourcePoint32 **parts, size_t nparts, const long *npts,
                                      const float *data, size_t nd, const float *h, size_t nh, size_t nm, long fields) {
    if (!set_sources(p, parts, nparts, npts, data, nd, fields)) return false;
    return set_interp(p, (long)nd, h, (long)nh, (long)nm, p->src_h, p->src_nh, p->src_nm);
}

// receivers with a coarse trace buffer (nr x nd, zeroed): each extraction at fine step it adds
// the reference's interpforward contributions of the fine sample (src/spacetime.cc:64-127)
static inline bool set_receivers_interp(Prop *p, SourcePoint32 **parts, size_t nparts, const long *npts, size_t nd,
                                        const float *h, size_t nh, size_t nm, long fields) {
    if (!set_receivers(p, parts, nparts, npts, nd, fields)) return false;
    return set_interp(p, (long)nd, h, (long)nh, (long)nm, p->rcv_h, p->rcv_nh, p->rcv_nm);
}

static inline bool get_receiver_data(Prop *p, float *out) {
    if (!p->rcv_data) return true;
    const size_t bytes = sizeof(float) * (size_t)p->rcv_nr * p->rcv_nt;
    WFD_TRY(hipMemcpyAsync(out, p->rcv_data, bytes, is_device_ptr(out) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                           p->stream));
    WFD_TRY(hipStreamSynchronize(p->stream));
    return true;
}

static inline int field_by_name(const char *name) {
    for (int i = 0; i < F_COUNT; i++)
        if (std::strcmp(name, FIELD_NAMES[i]) == 0) return i;
    return -1;
}

}  // namespace wfd
