// C-ABI layer: argument checks, the @SDF3 handle and the wrapper-level
// iteration loops.  Every loop issues its stage kernels back to back on the
// caller's stream; nothing here synchronizes except the calls that return a
// host scalar.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/vesicle_ls.h"
#include "vls_kernels.hpp"

using namespace vls;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define VLS_HIP(call)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) return fail(VLS_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                        \
    } while (0)

inline hipStream_t S(vls_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

// the error text of vls_last_error() for the host-only sources (vls_matio.cpp)
namespace vls {
int api_fail(int code, const char* msg) { return fail(code, "%s", msg); }
}  // namespace vls

namespace {

int check_geom(int rows, int cols, int pges, double dx, double dy, double dz)
{
    if (rows <= 0 || cols <= 0 || pges <= 0) return fail(VLS_ERR_INVALID, "box %dx%dx%d is empty", rows, cols, pges);
    if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return fail(VLS_ERR_INVALID, "spacings must be positive");
    if ((int64_t)rows * cols * pges > ((int64_t)1 << 40)) return fail(VLS_ERR_INVALID, "box too large");
    // the stage kernels address 7 pages with 32-bit byte offsets
    if ((int64_t)rows * cols >= ((int64_t)1 << 26)) return fail(VLS_ERR_INVALID, "a page holds 2^26 points or more");
    if (pges > 65535 || (cols + 3) / 4 > 65535) return fail(VLS_ERR_INVALID, "box exceeds the launch grid");
    return VLS_OK;
}

// compute units of the current device (256 on MI355X)
int cu_count()
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    return cus;
}

int check_num(int num_ele, int rows, int cols, int pges)
{
    if ((int64_t)num_ele != (int64_t)rows * cols * pges)
        return fail(VLS_ERR_INVALID, "num_ele %d != rows*cols*pges %lld", num_ele, (long long)rows * cols * pges);
    return VLS_OK;
}

template <typename... P>
bool any_null(P... p)
{
    return ((p == nullptr) || ...);
}

#define VLS_REQUIRE(cond, ...) \
    do {                       \
        if (!(cond)) return fail(VLS_ERR_INVALID, __VA_ARGS__); \
    } while (0)

Dist6 mk6(double* a, double* b, double* c, double* d, double* e, double* f)
{
    // argument order of the reference: xpr, xpl, ypf, ypb, zpu, zpd
    return Dist6{{a, c, e}, {b, d, f}};
}
Dist6c mk6c(const double* a, const double* b, const double* c, const double* d, const double* e, const double* f)
{
    return Dist6c{{a, c, e}, {b, d, f}};
}

// ---------------------------------------------------------------------------
// captured graphs.  A wrapper-level call (a whole reinitialization or
// extension loop, a surface redistance, a batch step) is captured once into a
// hipGraph on a private stream and replayed on the caller's stream after
// that.  The graph holds the call's device pointers and counts, so the cache
// is keyed by them; it keeps the 8 most recently used graphs per handle.  The
// captured work has no memset nodes: counters are cleared by kernels.
// ---------------------------------------------------------------------------
int g_graphs = 1;   // vls_set_graphs

struct GraphKey {
    uint64_t v[16] = {};
    bool operator==(const GraphKey& o) const { return std::memcmp(v, o.v, sizeof(v)) == 0; }
};

template <typename A>
uint64_t key_word(A a)
{
    if constexpr (std::is_pointer_v<A>) return (uint64_t)(uintptr_t)a;
    else if constexpr (std::is_floating_point_v<A>) {
        const double d = (double)a;
        uint64_t u;
        std::memcpy(&u, &d, sizeof(u));
        return u;
    } else return (uint64_t)(int64_t)a;
}

template <typename... A>
GraphKey make_key(A... a)
{
    static_assert(sizeof...(A) <= 16, "graph key too long");
    GraphKey k;
    int i = 0;
    ((k.v[i++] = key_word(a)), ...);
    return k;
}

struct GraphCache {
    struct Entry {
        GraphKey key;
        hipGraphExec_t exec;
        uint64_t used;
    };
    std::vector<Entry> entries;
    hipStream_t cap = nullptr;
    uint64_t tick = 0;
    void clear()
    {
        for (auto& e : entries) (void)hipGraphExecDestroy(e.exec);
        entries.clear();
    }
    ~GraphCache()
    {
        clear();
        if (cap) (void)hipStreamDestroy(cap);
    }
};

// run `issue(stream)` through the cache: replay the graph of `key`, or
// capture issue's launches into one and replay it
template <class Issue>
int with_graph(GraphCache& gc, const GraphKey& key0, hipStream_t s, Issue&& issue)
{
    if (!g_graphs) return issue(s);
    GraphKey key = key0;
    for (auto& e : gc.entries)
        if (e.key == key) {
            e.used = ++gc.tick;
            VLS_HIP(hipGraphLaunch(e.exec, s));
            return VLS_OK;
        }
    if (!gc.cap) VLS_HIP(hipStreamCreateWithFlags(&gc.cap, hipStreamNonBlocking));
    VLS_HIP(hipStreamBeginCapture(gc.cap, hipStreamCaptureModeThreadLocal));
    const int rc = issue(gc.cap);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(gc.cap, &graph);
    if (rc != VLS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(VLS_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(VLS_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    if (gc.entries.size() >= 8) {
        size_t lru = 0;
        for (size_t i = 1; i < gc.entries.size(); ++i)
            if (gc.entries[i].used < gc.entries[lru].used) lru = i;
        (void)hipGraphExecDestroy(gc.entries[lru].exec);
        gc.entries.erase(gc.entries.begin() + (long)lru);
    }
    gc.entries.push_back({key, exec, ++gc.tick});
    VLS_HIP(hipGraphLaunch(exec, s));
    return VLS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// the @SDF3 handle
// ---------------------------------------------------------------------------
struct vls_sdf3 {
    Geom g;
    double* pool = nullptr;  // every field below lives in one allocation
    uint8_t* flag_pool = nullptr;
    int* int_pool = nullptr;
    Dist6 d{};        // distances of the last set_distance (xpr..zpd)
    double* min_dist = nullptr;
    uint8_t* flags = nullptr;  // FLAG_SUBCELL of d, FLAG_MASK of the last reinit
    int* list = nullptr;       // points of d with a subcell distance, count in counts[0]
    Dist6 ad{};       // distances of an auxiliary field (ENO2 reinit, redistance of A)
    uint8_t* aflags = nullptr;
    int* alist = nullptr;      // counts[1]
    Dist6 fd{};       // distances of F inside the surface redistance
    uint8_t* fflags = nullptr;
    int* flist = nullptr;      // counts[2]
    int* over = nullptr;       // reinitialization points deferred by the bulk pass
    int* counts = nullptr;     // [0..2] list counts, [4..5] deferred counts
    double* ldist[3] = {nullptr, nullptr, nullptr};  // list-order distances of d, ad, fd (6 rows of n)
    double* ldt[3] = {nullptr, nullptr, nullptr};    // list-order pseudo-time steps
    double* lcp = nullptr;   // interface values of the extended field, list order (6 rows of n)
    double* lvel = nullptr;  // extension velocity at the listed points (3 rows of n)
    double* v[3] = {nullptr, nullptr, nullptr};
    double* t[4] = {nullptr, nullptr, nullptr, nullptr};
    double* partial = nullptr;
    double* result = nullptr;
    GraphCache graphs;
    bool dist_ready = false;  // a set_distance has filled flags and the list
    int fp64_tail = 0;        // FP64 iterations that end an FP32 reinitialization call
    bool dense_ready = false;  // d and min_dist hold the last set_distance's values (else built on demand)
    double* adj = nullptr;  // adjoint/tangent workspace, grown on demand
    size_t adj_len = 0;
};

namespace {

constexpr int kFields = 6 /*d*/ + 1 /*min*/ + 6 /*ad*/ + 6 /*fd*/ + 3 /*v*/ + 4 /*t*/ + 3 * 7 /*ldist, ldt*/ +
                        6 /*lcp*/ + 3 /*lvel*/;

Dist6 carve6(double*& p, int64_t n)
{
    Dist6 d;
    for (int a = 0; a < 3; ++a) { d.r[a] = p; p += n; d.l[a] = p; p += n; }
    return d;
}

// a distance set: the six arrays, their flags and their interface list
struct DSet {
    Dist6 d;
    uint8_t* flags;
    ListView lv;
};
DSet make_set(vls_sdf3* h, const Dist6& d, uint8_t* flags, int* list, int k)
{
    return DSet{d, flags, ListView{list, h->counts + k, h->ldist[k], h->ldt[k], (int)h->g.n()}};
}
DSet main_set(vls_sdf3* h) { return make_set(h, h->d, h->flags, h->list, 0); }
DSet aux_set(vls_sdf3* h) { return make_set(h, h->ad, h->aflags, h->alist, 1); }
DSet f_set(vls_sdf3* h) { return make_set(h, h->fd, h->fflags, h->flist, 2); }

// work fields of the handle as fields of T (FP32 calls use the first half of
// each FP64 field)
template <typename T>
T* tmp(vls_sdf3* h, int k)
{
    return reinterpret_cast<T*>(h->t[k]);
}
template <typename T>
T* vel(vls_sdf3* h, int a)
{
    return reinterpret_cast<T*>(h->v[a]);
}

// extension loop (WENO5RK3Extend.m:50-74, and the local extension of the
// surface redistance, WENO5RK3ClosetPointSurfaceRedistance.m:119-157):
// interface values of C once, then `iterations` RK3 steps; C in place.  The
// velocity (h->v) must be set; its listed values are gathered here.
template <typename T>
int extend_loop(vls_sdf3* h, T* C, int iterations, const DSet& ds, hipStream_t s)
{
    const Geom& g = h->g;
    T* t0 = tmp<T>(h, 0);
    T* t1 = tmp<T>(h, 1);
    const T *vx = vel<T>(h, 0), *vy = vel<T>(h, 1), *vz = vel<T>(h, 2);
    VLS_HIP(launch_list_interp(h->lcp, ds.lv, (const T*)C, g, s));
    for (int it = 0; it < iterations; ++it) {
        VLS_HIP(launch_extend_stage<T>(t0, C, nullptr, vx, vy, vz, ds.lv, h->lcp, h->lvel, g, 0, s));
        VLS_HIP(launch_extend_stage<T>(t1, t0, C, vx, vy, vz, ds.lv, h->lcp, h->lvel, g, 1, s));
        VLS_HIP(launch_extend_stage<T>(C, t1, C, vx, vy, vz, ds.lv, h->lcp, h->lvel, g, 2, s));
    }
    return VLS_OK;
}

// WENO5-RK3 reinitialization stages of one call share a sequence number for
// the deferred-point counters (see launch_reinit_stage)
struct ReinitSeq {
    int* counts;
    int seq;
};

template <typename T>
int reinit_stage(vls_sdf3* h, ReinitSeq& rs, T* out, const T* fs, const T* fn, const DSet& ds, int stage,
                 hipStream_t s)
{
    VLS_HIP(launch_reinit_stage<T>(out, fs, fn, ds.flags, ds.lv, h->g, stage, h->over, rs.counts, rs.seq++, s));
    return VLS_OK;
}

int reinit_begin(vls_sdf3* h, ReinitSeq& rs, hipStream_t s)
{
    rs.counts = h->counts + 4;
    rs.seq = 0;
    VLS_HIP(launch_zero_ints(rs.counts, 2, s));
    return VLS_OK;
}

// WENO5-RK3 reinitialization loop of F in place with mask and distances of `ds`
template <typename T>
int reinit_loop(vls_sdf3* h, T* F, int iterations, const DSet& ds, hipStream_t s)
{
    ReinitSeq rs;
    T* t0 = tmp<T>(h, 0);
    T* t1 = tmp<T>(h, 1);
    int rc = reinit_begin(h, rs, s);
    for (int it = 0; it < iterations && rc == VLS_OK; ++it) {
        if ((rc = reinit_stage<T>(h, rs, t0, F, nullptr, ds, 0, s))) break;
        if ((rc = reinit_stage<T>(h, rs, t1, t0, F, ds, 1, s))) break;
        rc = reinit_stage<T>(h, rs, F, t1, F, ds, 2, s);
    }
    return rc;
}

// the distances of F into the set `ds`: flags and interface list always;
// the dense arrays (six distances, and min_dist if given) only if `dense`
// (the kernels of the WENO5 path read the list; the ENO2 and cubic variants
// read the dense arrays).  The candidate points go through h->over and
// h->counts[3], free outside a reinitialization call.
template <typename T>
int set_distance_into(vls_sdf3* h, int order, const DSet& ds, double* min_dist, const T* F, const T* mask_src,
                      bool dense, hipStream_t s)
{
    const Dist6 d = dense ? ds.d : Dist6{};
    VLS_HIP(launch_set_distance<T>(order, d, dense ? min_dist : nullptr, ds.flags, mask_src, F, ds.lv, h->over,
                                   h->counts + 3, h->g, s));
    return VLS_OK;
}

// The last `tail` iterations of an FP32 reinitialization call in FP64: F is
// widened into the fourth work field (an FP64 field of the handle), iterated
// with the call's mask, flags and interface list (so the call is the same
// algorithm as one call of `iterations` iterations), and rounded back.
inline int reinit_tail_fp64(vls_sdf3* h, float* F, int tail, hipStream_t s)
{
    double* Fd = h->t[3];
    VLS_HIP(launch_widen(Fd, F, h->g.n(), s));
    const int rc = reinit_loop<double>(h, Fd, tail, main_set(h), s);
    if (rc != VLS_OK) return rc;
    VLS_HIP(launch_narrow(F, Fd, h->g.n(), s));
    return VLS_OK;
}
inline int reinit_tail_fp64(vls_sdf3*, double*, int, hipStream_t) { return VLS_OK; }

// ---- wrapper-level operations, for fields of T ------------------------------
template <typename T>
int sdf3_weno5_rk3_reinit(vls_sdf3* h, T* F, int iterations, hipStream_t s)
{
    VLS_REQUIRE(h && F, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    VLS_REQUIRE(h->dist_ready, "set_distance must run before the reinitialization (it reads its distances)");
    // an FP32 call may end with FP64 iterations (vls_sdf3_set_reinit_fp64_tail)
    const int tail = sizeof(T) == sizeof(float) ? std::min(h->fp64_tail, iterations) : 0;
    return with_graph(h->graphs, make_key(1, sizeof(T), F, iterations, tail), s, [&](hipStream_t cs) -> int {
        VLS_HIP(launch_set_mask_bit<T>(h->flags, F, h->g, cs));
        int rc = reinit_loop<T>(h, F, iterations - tail, main_set(h), cs);
        if (rc == VLS_OK && tail > 0) rc = reinit_tail_fp64(h, F, tail, cs);
        return rc;
    });
}

template <typename T>
int sdf3_weno5_rk3_extend(vls_sdf3* h, const T* F, T* C, int iterations, hipStream_t s)
{
    VLS_REQUIRE(h && F && C, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    VLS_REQUIRE(h->dist_ready, "set_distance must run before the extension (it reads F's distances)");
    VLS_REQUIRE((const void*)F != (const void*)C, "C must not alias F");
    const Geom& g = h->g;
    const DSet ms = main_set(h);
    T *vx = vel<T>(h, 0), *vy = vel<T>(h, 1), *vz = vel<T>(h, 2);
    return with_graph(h->graphs, make_key(2, sizeof(T), F, C, iterations), s, [&](hipStream_t cs) -> int {
        VLS_HIP(launch_extension_normal<T>(vx, vy, vz, F, h->flags, ms.lv, 1e-6, g, cs));
        VLS_HIP(launch_list_gather_vel<T>(h->lvel, ms.lv, vx, vy, vz, g, cs));
        return extend_loop<T>(h, C, iterations, ms, cs);
    });
}

template <typename T>
int surface_redistance_issue(vls_sdf3* h, const T* F, T* A, int iter1, int iter2, hipStream_t s);

template <typename T>
int sdf3_surface_redistance(vls_sdf3* h, const T* F, T* A, int iter1, int iter2, hipStream_t s)
{
    VLS_REQUIRE(h && F && A, "null argument");
    VLS_REQUIRE(iter1 >= 0 && iter2 >= 0, "negative iteration count");
    VLS_REQUIRE((const void*)F != (const void*)A, "A must not alias F");
    return with_graph(h->graphs, make_key(3, sizeof(T), F, A, iter1, iter2), s,
                      [&](hipStream_t cs) -> int { return surface_redistance_issue<T>(h, F, A, iter1, iter2, cs); });
}

template <typename T>
int surface_redistance_issue(vls_sdf3* h, const T* F, T* A, int iter1, int iter2, hipStream_t s)
{
    const Geom& g = h->g;
    int rc;
    // A's distances, deltat and mask (WENO5RK3ClosetPointSurfaceRedistance.m:8-25)
    const DSet as = aux_set(h), fs = f_set(h);
    if ((rc = set_distance_into<T>(h, 6, as, nullptr, A, A, false, s))) return rc;
    // F's distances and the extension velocity with floor 1e-14 (:30-74)
    if ((rc = set_distance_into<T>(h, 6, fs, nullptr, F, nullptr, false, s))) return rc;
    T *vx = vel<T>(h, 0), *vy = vel<T>(h, 1), *vz = vel<T>(h, 2);
    VLS_HIP(launch_extension_normal<T>(vx, vy, vz, F, fs.flags, fs.lv, 1e-14, g, s));
    VLS_HIP(launch_list_gather_vel<T>(h->lvel, fs.lv, vx, vy, vz, g, s));
    T* a1 = tmp<T>(h, 2);
    T* a05 = tmp<T>(h, 3);
    ReinitSeq rs;
    if ((rc = reinit_begin(h, rs, s))) return rc;
    for (int it = 0; it < iter1; ++it) {
        if ((rc = extend_loop<T>(h, A, iter2, fs, s))) return rc;
        if ((rc = reinit_stage<T>(h, rs, a1, A, nullptr, as, 0, s))) return rc;
        if ((rc = extend_loop<T>(h, a1, iter2, fs, s))) return rc;
        if ((rc = reinit_stage<T>(h, rs, a05, a1, A, as, 1, s))) return rc;
        if ((rc = extend_loop<T>(h, a05, iter2, fs, s))) return rc;
        if ((rc = reinit_stage<T>(h, rs, A, a05, A, as, 2, s))) return rc;
    }
    return extend_loop<T>(h, A, iter2, fs, s);
}

template <typename T>
int sdf3_advect_rk3(vls_sdf3* h, T* F, const T* vx, const T* vy, const T* vz, double dt, int steps, hipStream_t s)
{
    VLS_REQUIRE(h && F && vx && vy && vz, "null argument");
    VLS_REQUIRE(steps >= 0, "negative step count");
    const Geom& g = h->g;
    T* t0 = tmp<T>(h, 0);
    T* t1 = tmp<T>(h, 1);
    return with_graph(h->graphs, make_key(4, sizeof(T), F, vx, vy, vz, dt, steps), s, [&](hipStream_t cs) -> int {
        for (int it = 0; it < steps; ++it) {
            VLS_HIP(launch_advect_stage<T>(t0, F, nullptr, vx, vy, vz, dt, g, 0, cs));
            VLS_HIP(launch_advect_stage<T>(t1, t0, F, vx, vy, vz, dt, g, 1, cs));
            VLS_HIP(launch_advect_stage<T>(F, t1, F, vx, vy, vz, dt, g, 2, cs));
        }
        return VLS_OK;
    });
}

template <typename T>
int sdf3_max_speed(vls_sdf3* h, const T* vx, const T* vy, const T* vz, double* out, hipStream_t s);

template <typename T>
int area_volume(vls_sdf3* h, int kind, const T* F, double eps, double* dev_out, double* area, double* volume,
                hipStream_t s);

template <typename T>
int sdf3_normal_velocity(vls_sdf3* h, const T* F, const T* vx, const T* vy, const T* vz, T* out, hipStream_t s)
{
    VLS_REQUIRE(h && F && vx && vy && vz && out, "null argument");
    VLS_REQUIRE((const void*)out != (const void*)F, "out must not alias F");
    VLS_HIP(launch_normal_velocity<T>(out, F, vx, vy, vz, h->g, s));
    return VLS_OK;
}

int read_result(vls_sdf3* h, double* out, int n, hipStream_t s)
{
    VLS_HIP(hipMemcpyAsync(out, h->result, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    VLS_HIP(hipStreamSynchronize(s));
    return VLS_OK;
}

template <typename T>
int sdf3_max_speed(vls_sdf3* h, const T* vx, const T* vy, const T* vz, double* out, hipStream_t s)
{
    VLS_REQUIRE(h && vx && vy && vz && out, "null argument");
    VLS_HIP(launch_max_abs3<T>(h->result, h->partial, vx, vy, vz, h->g, s));
    return read_result(h, out, 1, s);
}

template <typename T>
int area_volume(vls_sdf3* h, int kind, const T* F, double eps, double* dev_out, double* area, double* volume,
                hipStream_t s)
{
    VLS_REQUIRE(h && F, "null argument");
    VLS_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (Towers) or 1 (cosine)");
    const Geom& g = h->g;
    double* dst = dev_out ? dev_out : h->result;
    if (kind == 0) {
        const double ds = (g.dx + g.dy + g.dz) / 3;
        VLS_HIP(launch_area_volume_towers<T>(dst, h->partial, F, g, ds, ds * ds * ds, s));
    } else {
        VLS_REQUIRE(eps > 0, "eps must be positive");
        VLS_HIP(launch_area_volume_cosine<T>(dst, h->partial, F, g, eps, g.dx * g.dy * g.dz, s));
    }
    if (dev_out) return VLS_OK;
    double r[2];
    int rc = read_result(h, r, 2, s);
    if (rc) return rc;
    *area = r[0];
    *volume = r[1];
    return VLS_OK;
}

// ---- work memory of the allocation-free reinitialization --------------------
// One copy per GPU, part of the library's code object: the runtime places it
// when it loads the code object on a device and it stays until the process
// ends.  Per grid point: the list-order distances (six rows) and
// pseudo-time steps, two RK stage fields (9 doubles), the interface list and
// the deferred points (2 ints) and the flags (1 byte); 16 counters; every
// piece starts on a 256-byte boundary.
constexpr int kStaticPieces = 8;
constexpr size_t kStaticBytes =
    (size_t)VLS_STATIC_POINTS * (9 * sizeof(double) + 2 * sizeof(int) + 1) + 16 * sizeof(int) + kStaticPieces * 256;
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// the device address of the current device's copy of `scratch`, looked up
// once per device (the array's name is a host-side stand-in in host code)
int static_scratch(unsigned char** out)
{
    int dev = 0;
    VLS_HIP(hipGetDevice(&dev));
    VLS_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(VLS_ERR_HIP, "hipGetSymbolAddress of the static work memory on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return VLS_OK;
}

}  // namespace

extern "C" {

const char* vls_last_error(void) { return g_err.c_str(); }

int vls_set_graphs(int enable)
{
    const int prev = g_graphs;
    g_graphs = enable ? 1 : 0;
    return prev;
}

const char* vls_version(void) { return "vesicle_ls 0.1 (gfx950, FP64)"; }

// ---------------------------------------------------------------------------
// 1. kernel-level drop-ins
// ---------------------------------------------------------------------------
static int distances_dropin(int order, double* xpr, double* xpl, double* ypf, double* ypb, double* zpu,
                            double* zpd, const double* lsf, int num_ele, int rows, int cols, int pges, double dx,
                            double dy, double dz, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(xpr, xpl, ypf, ypb, zpu, zpd, lsf), "null pointer");
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_boundary_location(order, mk6(xpr, xpl, ypf, ypb, zpu, zpd), lsf, g, S(stream)));
    return VLS_OK;
}

int vls_boundary_location(double* xpr, double* xpl, double* ypf, double* ypb, double* zpu, double* zpd,
                          const double* lsf, int num_ele, int rows, int cols, int pges, double dx, double dy,
                          double dz, vls_stream_t stream)
{
    return distances_dropin(6, xpr, xpl, ypf, ypb, zpu, zpd, lsf, num_ele, rows, cols, pges, dx, dy, dz, stream);
}

int vls_boundary_correction(double* xpr, double* xpl, double* ypf, double* ypb, double* zpu, double* zpd,
                            const double* lsf, int num_ele, int rows, int cols, int pges, double dx, double dy,
                            double dz, vls_stream_t stream)
{
    return distances_dropin(2, xpr, xpl, ypf, ypb, zpu, zpd, lsf, num_ele, rows, cols, pges, dx, dy, dz, stream);
}

static int re_step_dropin(int eno2, double* step, const double* lsf, const uint8_t* mask, const double* deltat,
                          const double* xpr, const double* xpl, const double* ypf, const double* ypb,
                          const double* zpu, const double* zpd, int rows, int cols, int pges, double dx,
                          double dy, double dz, int num_ele, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(step, lsf, mask, deltat, xpr, xpl, ypf, ypb, zpu, zpd), "null pointer");
    VLS_REQUIRE(step != lsf, "step must not alias lsf");
    const Geom g{rows, cols, pges, dx, dy, dz};
    const Dist6c d = mk6c(xpr, xpl, ypf, ypb, zpu, zpd);
    if (eno2) VLS_HIP(launch_eno2_re_step(step, lsf, mask, deltat, d, g, S(stream)));
    else VLS_HIP(launch_re_step(step, lsf, mask, deltat, d, g, S(stream)));
    return VLS_OK;
}

int vls_re_step(double* step, const double* lsf, const uint8_t* mask, const double* deltat, const double* xpr,
                const double* xpl, const double* ypf, const double* ypb, const double* zpu, const double* zpd,
                int rows, int cols, int pges, double dx, double dy, double dz, int num_ele, vls_stream_t stream)
{
    return re_step_dropin(0, step, lsf, mask, deltat, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz,
                          num_ele, stream);
}

int vls_eno2_re_step(double* step, const double* lsf, const uint8_t* mask, const double* deltat,
                     const double* xpr, const double* xpl, const double* ypf, const double* ypb,
                     const double* zpu, const double* zpd, int rows, int cols, int pges, double dx, double dy,
                     double dz, int num_ele, vls_stream_t stream)
{
    return re_step_dropin(1, step, lsf, mask, deltat, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz,
                          num_ele, stream);
}

int vls_upwind_normal(double* nx, double* ny, double* nz, const double* lsf, const double* xpr, const double* xpl,
                      const double* ypf, const double* ypb, const double* zpu, const double* zpd, int rows,
                      int cols, int pges, double dx, double dy, double dz, int num_ele, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(nx, ny, nz, lsf, xpr, xpl, ypf, ypb, zpu, zpd), "null pointer");
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_upwind_normal(nx, ny, nz, lsf, mk6c(xpr, xpl, ypf, ypb, zpu, zpd), g, S(stream)));
    return VLS_OK;
}

int vls_boundary_interpolant(double* cpr, double* cpl, double* cpf, double* cpb, double* cpu, double* cpd,
                             const double* xpr, const double* xpl, const double* ypf, const double* ypb,
                             const double* zpu, const double* zpd, const double* lsf, int rows, int cols,
                             int pges, double dx, double dy, double dz, int num_ele, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(cpr, cpl, cpf, cpb, cpu, cpd, xpr, xpl, ypf, ypb, zpu, zpd, lsf), "null pointer");
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_boundary_interpolant(mk6(cpr, cpl, cpf, cpb, cpu, cpd), mk6c(xpr, xpl, ypf, ypb, zpu, zpd), lsf,
                                        g, S(stream)));
    return VLS_OK;
}

int vls_extend_step(double* step, const double* deltat, const double* lsf, const double* vx, const double* vy,
                    const double* vz, const double* xpr, const double* xpl, const double* ypf, const double* ypb,
                    const double* zpu, const double* zpd, const double* cpr, const double* cpl, const double* cpf,
                    const double* cpb, const double* cpu, const double* cpd, int rows, int cols, int pges,
                    double dx, double dy, double dz, int num_ele, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(step, deltat, lsf, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cpr, cpl, cpf, cpb, cpu, cpd),
                "null pointer");
    VLS_REQUIRE(step != lsf, "step must not alias lsf");
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_extend_step(step, deltat, lsf, vx, vy, vz, mk6c(xpr, xpl, ypf, ypb, zpu, zpd),
                               mk6c(cpr, cpl, cpf, cpb, cpu, cpd), g, S(stream)));
    return VLS_OK;
}

// ---- WENO-RK3 / ENO-RK2 variants: drop-in kernels -----------------------------
int vls_cubic_boundary_correction(double* xpr, double* xpl, double* ypf, double* ypb, double* zpu, double* zpd,
                                  const double* lsf, int num_ele, int rows, int cols, int pges, double dx,
                                  double dy, double dz, vls_stream_t stream)
{
    return distances_dropin(3, xpr, xpl, ypf, ypb, zpu, zpd, lsf, num_ele, rows, cols, pges, dx, dy, dz, stream);
}

int vls_wenork3_re_step(double* step, const double* lsf, const uint8_t* mask, const double* deltat,
                        const double* xpr, const double* xpl, const double* ypf, const double* ypb,
                        const double* zpu, const double* zpd, int rows, int cols, int pges, double dx, double dy,
                        double dz, int num_ele, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(step, lsf, mask, deltat, xpr, xpl, ypf, ypb, zpu, zpd), "null pointer");
    VLS_REQUIRE(step != lsf, "step must not alias lsf");
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_wenork3_re_step(step, lsf, mask, deltat, mk6c(xpr, xpl, ypf, ypb, zpu, zpd), g, S(stream)));
    return VLS_OK;
}

static int variant_normal_dropin(int weno, double* nx, double* ny, double* nz, const double* lsf, const double* xpr,
                                 const double* xpl, const double* ypf, const double* ypb, const double* zpu,
                                 const double* zpd, int rows, int cols, int pges, double dx, double dy, double dz,
                                 int num_ele, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(nx, ny, nz, lsf, xpr, xpl, ypf, ypb, zpu, zpd), "null pointer");
    const Geom g{rows, cols, pges, dx, dy, dz};
    const Dist6c d = mk6c(xpr, xpl, ypf, ypb, zpu, zpd);
    if (weno) VLS_HIP(launch_wenork3_upwind_normal(nx, ny, nz, lsf, d, g, S(stream)));
    else VLS_HIP(launch_enork2_upwind_normal(nx, ny, nz, lsf, d, g, S(stream)));
    return VLS_OK;
}

int vls_enork2_upwind_normal(double* nx, double* ny, double* nz, const double* lsf, const double* xpr,
                             const double* xpl, const double* ypf, const double* ypb, const double* zpu,
                             const double* zpd, int rows, int cols, int pges, double dx, double dy, double dz,
                             int num_ele, vls_stream_t stream)
{
    return variant_normal_dropin(0, nx, ny, nz, lsf, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz,
                                 num_ele, stream);
}

int vls_wenork3_upwind_normal(double* nx, double* ny, double* nz, const double* lsf, const double* xpr,
                              const double* xpl, const double* ypf, const double* ypb, const double* zpu,
                              const double* zpd, int rows, int cols, int pges, double dx, double dy, double dz,
                              int num_ele, vls_stream_t stream)
{
    return variant_normal_dropin(1, nx, ny, nz, lsf, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz,
                                 num_ele, stream);
}

static int variant_interpolate_dropin(int weno, double* cpr, double* cpl, double* cpf, double* cpb, double* cpu,
                                      double* cpd, const double* xpr, const double* xpl, const double* ypf,
                                      const double* ypb, const double* zpu, const double* zpd, const double* lsf,
                                      int rows, int cols, int pges, double dx, double dy, double dz, int num_ele,
                                      vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(cpr, cpl, cpf, cpb, cpu, cpd, xpr, xpl, ypf, ypb, zpu, zpd, lsf), "null pointer");
    const Geom g{rows, cols, pges, dx, dy, dz};
    const Dist6 cp = mk6(cpr, cpl, cpf, cpb, cpu, cpd);
    const Dist6c d = mk6c(xpr, xpl, ypf, ypb, zpu, zpd);
    if (weno) VLS_HIP(launch_wenork3_boundary_interpolate(cp, d, lsf, g, S(stream)));
    else VLS_HIP(launch_enork2_boundary_interpolate(cp, d, lsf, g, S(stream)));
    return VLS_OK;
}

int vls_enork2_boundary_interpolate(double* cpr, double* cpl, double* cpf, double* cpb, double* cpu, double* cpd,
                                    const double* xpr, const double* xpl, const double* ypf, const double* ypb,
                                    const double* zpu, const double* zpd, const double* lsf, int rows, int cols,
                                    int pges, double dx, double dy, double dz, int num_ele, vls_stream_t stream)
{
    return variant_interpolate_dropin(0, cpr, cpl, cpf, cpb, cpu, cpd, xpr, xpl, ypf, ypb, zpu, zpd, lsf, rows,
                                      cols, pges, dx, dy, dz, num_ele, stream);
}

int vls_wenork3_boundary_interpolate(double* cpr, double* cpl, double* cpf, double* cpb, double* cpu, double* cpd,
                                     const double* xpr, const double* xpl, const double* ypf, const double* ypb,
                                     const double* zpu, const double* zpd, const double* lsf, int rows, int cols,
                                     int pges, double dx, double dy, double dz, int num_ele, vls_stream_t stream)
{
    return variant_interpolate_dropin(1, cpr, cpl, cpf, cpb, cpu, cpd, xpr, xpl, ypf, ypb, zpu, zpd, lsf, rows,
                                      cols, pges, dx, dy, dz, num_ele, stream);
}

static int variant_extend_dropin(int weno, double* step, const double* deltat, const double* lsf, const double* vx,
                                 const double* vy, const double* vz, const double* xpr, const double* xpl,
                                 const double* ypf, const double* ypb, const double* zpu, const double* zpd,
                                 const double* cpr, const double* cpl, const double* cpf, const double* cpb,
                                 const double* cpu, const double* cpd, int rows, int cols, int pges, double dx,
                                 double dy, double dz, int num_ele, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(step, deltat, lsf, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cpr, cpl, cpf, cpb, cpu, cpd),
                "null pointer");
    VLS_REQUIRE(step != lsf, "step must not alias lsf");
    const Geom g{rows, cols, pges, dx, dy, dz};
    const Dist6c d = mk6c(xpr, xpl, ypf, ypb, zpu, zpd), cp = mk6c(cpr, cpl, cpf, cpb, cpu, cpd);
    if (weno) VLS_HIP(launch_wenork3_extend_step(step, deltat, lsf, vx, vy, vz, d, cp, g, S(stream)));
    else VLS_HIP(launch_enork2_extend_step(step, deltat, lsf, vx, vy, vz, d, cp, g, S(stream)));
    return VLS_OK;
}

int vls_enork2_extend_step(double* step, const double* deltat, const double* lsf, const double* vx, const double* vy,
                           const double* vz, const double* xpr, const double* xpl, const double* ypf,
                           const double* ypb, const double* zpu, const double* zpd, const double* cpr,
                           const double* cpl, const double* cpf, const double* cpb, const double* cpu,
                           const double* cpd, int rows, int cols, int pges, double dx, double dy, double dz,
                           int num_ele, vls_stream_t stream)
{
    return variant_extend_dropin(0, step, deltat, lsf, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cpr, cpl, cpf, cpb,
                                 cpu, cpd, rows, cols, pges, dx, dy, dz, num_ele, stream);
}

int vls_wenork3_extend_step(double* step, const double* deltat, const double* lsf, const double* vx,
                            const double* vy, const double* vz, const double* xpr, const double* xpl,
                            const double* ypf, const double* ypb, const double* zpu, const double* zpd,
                            const double* cpr, const double* cpl, const double* cpf, const double* cpb,
                            const double* cpu, const double* cpd, int rows, int cols, int pges, double dx, double dy,
                            double dz, int num_ele, vls_stream_t stream)
{
    return variant_extend_dropin(1, step, deltat, lsf, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cpr, cpl, cpf, cpb,
                                 cpu, cpd, rows, cols, pges, dx, dy, dz, num_ele, stream);
}

int vls_advection_step(double* astep, const double* vx, const double* vy, const double* vz, const double* lsf,
                       int rows, int cols, int pges, double dx, double dy, double dz, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    VLS_REQUIRE(!any_null(astep, vx, vy, vz, lsf), "null pointer");
    VLS_REQUIRE(astep != lsf, "astep must not alias lsf");
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_advection_step(astep, vx, vy, vz, lsf, g, S(stream)));
    return VLS_OK;
}

int vls_weno_derivative(double* back_x, double* fore_x, double* back_y, double* fore_y, double* back_z,
                        double* fore_z, const double* lsf, int rows, int cols, int pges, double dx, double dy,
                        double dz, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    VLS_REQUIRE(!any_null(back_x, fore_x, back_y, fore_y, back_z, fore_z, lsf), "null pointer");
    const Geom g{rows, cols, pges, dx, dy, dz};
    // Dist6 slots: r = fore, l = back
    VLS_HIP(launch_weno_derivative(Dist6{{fore_x, fore_y, fore_z}, {back_x, back_y, back_z}}, lsf, g, S(stream)));
    return VLS_OK;
}

int vls_set_calculus_toolbox(double* const* out18, const double* lsf, const double* hprimal, int rows, int cols,
                             int pges, double dx, double dy, double dz, double ds, int num_ele,
                             vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(out18, lsf, hprimal), "null pointer");
    Fields18 o;
    for (int k = 0; k < 18; ++k) {
        VLS_REQUIRE(out18[k] != nullptr, "null output %d", k);
        o.f[k] = out18[k];
    }
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_set_calculus_toolbox(o, lsf, hprimal, g, ds, S(stream)));
    return VLS_OK;
}

int vls_auxi_set_calculus_toolbox(double* const* out27, const double* lsf, const double* ahprimal,
                                  const double* const* fin13, int rows, int cols, int pges, double dx, double dy,
                                  double dz, double ds, int num_ele, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    if ((rc = check_num(num_ele, rows, cols, pges))) return rc;
    VLS_REQUIRE(!any_null(out27, lsf, ahprimal, fin13), "null pointer");
    Fields27 o;
    for (int k = 0; k < 27; ++k) {
        VLS_REQUIRE(out27[k] != nullptr, "null output %d", k);
        o.f[k] = out27[k];
    }
    Fields13c in;
    for (int k = 0; k < 13; ++k) {
        VLS_REQUIRE(fin13[k] != nullptr, "null input field %d", k);
        in.f[k] = fin13[k];
    }
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_auxi_set_calculus_toolbox(o, lsf, ahprimal, in, g, ds, S(stream)));
    return VLS_OK;
}

int vls_aux_calculus_toolbox(double* const* out30, const double* A, const double* const* ftool14, int rows,
                             int cols, int pges, double dx, double dy, double dz, double ds, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    VLS_REQUIRE(!any_null(out30, A, ftool14), "null pointer");
    for (int k = 0; k < 30; ++k) VLS_REQUIRE(out30[k] != nullptr, "null output %d", k);
    for (int k = 0; k < 14; ++k) VLS_REQUIRE(ftool14[k] != nullptr, "null input field %d", k);
    const Geom g{rows, cols, pges, dx, dy, dz};
    // out30 = {AHPrimal, the 27 kernel outputs, AFDiracDelta, ADiracDeltaDn}
    VLS_HIP(launch_positive_part(out30[0], A, g.n(), S(stream)));
    Fields27 o;
    for (int k = 0; k < 27; ++k) o.f[k] = out30[1 + k];
    Fields13c in;
    for (int k = 0; k < 13; ++k) in.f[k] = ftool14[k];
    VLS_HIP(launch_auxi_set_calculus_toolbox(o, A, out30[0], in, g, ds, S(stream)));
    // out30 positions: AGradMag 4, C 5..7, NormCrossAF 8, AHeaviside 26, ADiracDelta 27;
    // F's Heaviside is ftool14[13]
    VLS_HIP(launch_aux_toolbox_tail(out30[28], out30[29], out30[26], ftool14[13], A, out30[5], out30[6], out30[7],
                                    out30[8], out30[27], out30[4], g, S(stream)));
    return VLS_OK;
}

static int smooth_entry(double* out, const double* rhs, int rows, int cols, int pges, double dx, double dy,
                        double dz, double dt, double alpha, int diffusion, int exact, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    VLS_REQUIRE(!any_null(out, rhs), "null pointer");
    VLS_REQUIRE(diffusion == 0 || diffusion == 1, "diffusion must be 0 (bilaplacian) or 1 (laplacian)");
    VLS_REQUIRE(std::isfinite(dt) && std::isfinite(alpha), "dt and alpha must be finite");
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_smooth(out, rhs, g, alpha * dt, diffusion, exact, S(stream)));
    return VLS_OK;
}

int vls_smooth_fft(double* out, const double* rhs, int rows, int cols, int pges, double dx, double dy, double dz,
                   double dt, double alpha, int diffusion, vls_stream_t stream)
{
    return smooth_entry(out, rhs, rows, cols, pges, dx, dy, dz, dt, alpha, diffusion, 0, stream);
}

int vls_smooth_gmres(double* out, const double* rhs, int rows, int cols, int pges, double dx, double dy, double dz,
                     double dt, double alpha, int diffusion, vls_stream_t stream)
{
    return smooth_entry(out, rhs, rows, cols, pges, dx, dy, dz, dt, alpha, diffusion, 1, stream);
}

int vls_smooth_operator(double* y, const double* x, int rows, int cols, int pges, double dx, double dy, double dz,
                        double dt, double alpha, int diffusion, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    VLS_REQUIRE(!any_null(y, x), "null pointer");
    VLS_REQUIRE(y != x, "the operator cannot be applied in place");
    VLS_REQUIRE(diffusion == 0 || diffusion == 1, "diffusion must be 0 (bilaplacian) or 1 (laplacian)");
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_smooth_operator(y, x, g, alpha * dt, diffusion, S(stream)));
    return VLS_OK;
}

int vls_bending_speed(double* out, const double* H, const double* G, int rows, int cols, int pges, double dx,
                      double dy, double dz, double kappa, double sponc, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    VLS_REQUIRE(!any_null(out, H, G), "null pointer");
    VLS_REQUIRE(out != H, "the bending speed cannot overwrite H (its Laplacian reads neighbours)");
    const Geom g{rows, cols, pges, dx, dy, dz};
    VLS_HIP(launch_bending_speed(out, H, G, g, kappa, sponc, S(stream)));
    return VLS_OK;
}

// ---------------------------------------------------------------------------
// 2. @SDF3 handle
// ---------------------------------------------------------------------------
int vls_sdf3_create(vls_sdf3_t* out, int rows, int cols, int pges, double dx, double dy, double dz)
{
    VLS_REQUIRE(out != nullptr, "null handle pointer");
    *out = nullptr;
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    vls_sdf3* h = new (std::nothrow) vls_sdf3();
    if (!h) return fail(VLS_ERR_ALLOC, "host allocation failed");
    h->g = Geom{rows, cols, pges, dx, dy, dz};
    const int64_t n = h->g.n();
    const size_t bytes = sizeof(double) * ((size_t)n * kFields + (size_t)reduce_blocks() * 16 + 16);
    if (hipMalloc((void**)&h->pool, bytes) != hipSuccess) {
        delete h;
        return fail(VLS_ERR_ALLOC, "hipMalloc of %zu bytes failed", bytes);
    }
    if (hipMalloc((void**)&h->flag_pool, (size_t)n * 3) != hipSuccess) {
        (void)hipFree(h->pool);
        delete h;
        return fail(VLS_ERR_ALLOC, "hipMalloc of flags failed");
    }
    if (hipMalloc((void**)&h->int_pool, sizeof(int) * ((size_t)n * 4 + 16)) != hipSuccess) {
        (void)hipFree(h->pool);
        (void)hipFree(h->flag_pool);
        delete h;
        return fail(VLS_ERR_ALLOC, "hipMalloc of point lists failed");
    }
    if (hipMemset(h->int_pool, 0, sizeof(int) * ((size_t)n * 4 + 16)) != hipSuccess ||
        hipMemset(h->flag_pool, 0, (size_t)n * 3) != hipSuccess) {
        (void)hipFree(h->pool);
        (void)hipFree(h->flag_pool);
        (void)hipFree(h->int_pool);
        delete h;
        return fail(VLS_ERR_HIP, "hipMemset of point lists failed");
    }
    double* p = h->pool;
    h->d = carve6(p, n);
    h->min_dist = p; p += n;
    h->ad = carve6(p, n);
    h->fd = carve6(p, n);
    for (int k = 0; k < 3; ++k) {
        h->ldist[k] = p; p += 6 * n;
        h->ldt[k] = p; p += n;
    }
    h->lcp = p; p += 6 * n;
    h->lvel = p; p += 3 * n;
    for (int a = 0; a < 3; ++a) { h->v[a] = p; p += n; }
    for (int k = 0; k < 4; ++k) { h->t[k] = p; p += n; }
    h->partial = p; p += (size_t)reduce_blocks() * 16;
    h->result = p;
    h->flags = h->flag_pool;
    h->aflags = h->flag_pool + n;
    h->fflags = h->flag_pool + 2 * n;
    h->list = h->int_pool;
    h->alist = h->int_pool + n;
    h->flist = h->int_pool + 2 * n;
    h->over = h->int_pool + 3 * n;
    h->counts = h->int_pool + 4 * n;
    *out = h;
    return VLS_OK;
}

int vls_sdf3_set_tile_rows(vls_sdf3_t h, int tile_rows)
{
    VLS_REQUIRE(h, "null handle");
    VLS_REQUIRE(tile_rows == 32 || tile_rows == 64, "tile rows must be 32 or 64");
    h->g.tr = tile_rows;
    h->graphs.clear();
    return VLS_OK;
}

int vls_sdf3_set_zmarch_depth(vls_sdf3_t h, int depth)
{
    VLS_REQUIRE(h, "null handle");
    VLS_REQUIRE(depth >= 0, "depth must be 0 (by the box size) or positive");
    h->g.zd = depth;
    h->graphs.clear();
    return VLS_OK;
}

int vls_sdf3_get_launch(vls_sdf3_t h, int* tile_rows, int* depth)
{
    VLS_REQUIRE(h, "null handle");
    if (tile_rows) *tile_rows = h->g.tr;
    if (depth) *depth = vls::zmarch_depth(h->g);
    return VLS_OK;
}

int vls_sdf3_destroy(vls_sdf3_t h)
{
    if (!h) return VLS_OK;
    (void)hipFree(h->adj);
    hipError_t e1 = hipFree(h->pool), e2 = hipFree(h->flag_pool), e3 = hipFree(h->int_pool);
    delete h;
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(VLS_ERR_HIP, "hipFree failed");
    return VLS_OK;
}

int vls_sdf3_set_distance(vls_sdf3_t h, const double* F, vls_stream_t stream)
{
    VLS_REQUIRE(h && F, "null argument");
    const int rc = set_distance_into<double>(h, 6, main_set(h), h->min_dist, F, nullptr, false, S(stream));
    if (rc == VLS_OK) h->dist_ready = true;
    h->dense_ready = false;
    return rc;
}

int vls_sdf3_set_distance_f32(vls_sdf3_t h, const float* F, vls_stream_t stream)
{
    VLS_REQUIRE(h && F, "null argument");
    const int rc = set_distance_into<float>(h, 6, main_set(h), h->min_dist, F, nullptr, false, S(stream));
    if (rc == VLS_OK) h->dist_ready = true;
    h->dense_ready = false;
    return rc;
}

int vls_sdf3_distances(vls_sdf3_t h, const double** out7)
{
    VLS_REQUIRE(h && out7, "null argument");
    if (h->dist_ready && !h->dense_ready) {
        // the dense arrays of the last set_distance, from its list, after
        // all queued work, waited for: the pointers are then ready for any
        // stream
        VLS_HIP(hipDeviceSynchronize());
        VLS_HIP(launch_fill_dense(h->d, h->min_dist, main_set(h).lv, h->g, nullptr));
        VLS_HIP(hipStreamSynchronize(nullptr));
        h->dense_ready = true;
    }
    out7[0] = h->d.r[0]; out7[1] = h->d.l[0];
    out7[2] = h->d.r[1]; out7[3] = h->d.l[1];
    out7[4] = h->d.r[2]; out7[5] = h->d.l[2];
    out7[6] = h->min_dist;
    return VLS_OK;
}

int vls_sdf3_weno5_rk3_reinit(vls_sdf3_t h, double* F, int iterations, vls_stream_t stream)
{
    return sdf3_weno5_rk3_reinit<double>(h, F, iterations, S(stream));
}

int vls_sdf3_weno5_rk3_reinit_f32(vls_sdf3_t h, float* F, int iterations, vls_stream_t stream)
{
    return sdf3_weno5_rk3_reinit<float>(h, F, iterations, S(stream));
}

int vls_sdf3_set_reinit_fp64_tail(vls_sdf3_t h, int iterations)
{
    VLS_REQUIRE(h, "null handle");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    h->fp64_tail = iterations;
    return VLS_OK;
}

int vls_weno5_rk3_reinit_static(double* F, int iterations, int rows, int cols, int pges, double dx, double dy,
                                double dz, vls_stream_t stream)
{
    int rc = check_geom(rows, cols, pges, dx, dy, dz);
    if (rc) return rc;
    VLS_REQUIRE(F, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    const int64_t n = (int64_t)rows * cols * pges;
    VLS_REQUIRE(n <= VLS_STATIC_POINTS, "box %dx%dx%d has %lld points, more than the %d the static work memory holds",
                rows, cols, pges, (long long)n, VLS_STATIC_POINTS);
    unsigned char* base = nullptr;
    if ((rc = static_scratch(&base))) return rc;
    // the handle's fields this path uses, carved from the static memory
    vls_sdf3 h;
    h.g = Geom{rows, cols, pges, dx, dy, dz};
    size_t used = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + used;
        used += (bytes + 255) / 256 * 256;
        return p;
    };
    h.ldist[0] = reinterpret_cast<double*>(take(sizeof(double) * 6 * n));
    h.ldt[0] = reinterpret_cast<double*>(take(sizeof(double) * n));
    h.t[0] = reinterpret_cast<double*>(take(sizeof(double) * n));
    h.t[1] = reinterpret_cast<double*>(take(sizeof(double) * n));
    h.list = reinterpret_cast<int*>(take(sizeof(int) * n));
    h.over = reinterpret_cast<int*>(take(sizeof(int) * n));
    h.counts = reinterpret_cast<int*>(take(sizeof(int) * 16));
    h.flags = take((size_t)n);
    VLS_REQUIRE(used <= kStaticBytes, "static work memory layout overflows (%zu > %zu bytes)", used, kStaticBytes);
    const hipStream_t s = S(stream);
    // the counters keep the previous call's values: clear them (a kernel, so
    // that a captured call records no memset node)
    VLS_HIP(launch_zero_ints(h.counts, 16, s));
    if ((rc = set_distance_into<double>(&h, 6, main_set(&h), nullptr, F, nullptr, false, s))) return rc;
    VLS_HIP(launch_set_mask_bit<double>(h.flags, F, h.g, s));
    return reinit_loop<double>(&h, F, iterations, main_set(&h), s);
}

int vls_sdf3_eno2_reinit(vls_sdf3_t h, double* F, int iterations, int rk2, double cfl, vls_stream_t stream)
{
    VLS_REQUIRE(h && F, "null argument");
    VLS_REQUIRE(iterations >= 0 && cfl > 0, "bad iterations/cfl");
    const hipStream_t s = S(stream);
    const Geom& g = h->g;
    // ENO2 distances of the argument and mask = F<0 (ENORK2Reinitialization.m:7-25)
    int rc = set_distance_into<double>(h, 2, aux_set(h), nullptr, F, F, true, s);
    if (rc) return rc;
    const Dist6c dc = constify(h->ad);
    if (rk2) {
        for (int it = 0; it < iterations; ++it) {
            VLS_HIP(launch_eno2_stage(h->t[0], F, nullptr, h->aflags, dc, g, cfl, 0, s));
            VLS_HIP(launch_eno2_stage(F, h->t[0], F, h->aflags, dc, g, cfl, 1, s));
        }
        return VLS_OK;
    }
    // forward Euler: ping-pong between F and t[0]
    double* cur = F;
    double* nxt = h->t[0];
    for (int it = 0; it < iterations; ++it) {
        VLS_HIP(launch_eno2_stage(nxt, cur, nullptr, h->aflags, dc, g, cfl, 0, s));
        double* tmp = cur; cur = nxt; nxt = tmp;
    }
    if (cur != F) VLS_HIP(launch_copy<double>(F, cur, g.n(), s));
    return VLS_OK;
}

// WENORK3Reinitialization (+SD/@SDF3/WENORK3Reinitialization.m:3-61): cubic
// distances of the argument (:17-21) into the auxiliary distance set, mask =
// F<0, dt = 0.3*min (:23-26), RK3 (:30-57) on the fused stage kernels of the
// WENO5 path with cubic ENO at the interface
int vls_sdf3_wenork3_reinit(vls_sdf3_t h, double* F, int iterations, vls_stream_t stream)
{
    VLS_REQUIRE(h && F, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    return with_graph(h->graphs, make_key(21, F, iterations), S(stream), [&](hipStream_t s) -> int {
        const Geom& g = h->g;
        const DSet as = aux_set(h);
        int rc = set_distance_into<double>(h, 3, as, nullptr, F, F, false, s);
        if (rc) return rc;
        double *t0 = h->t[0], *t1 = h->t[1];
        ReinitSeq rs;
        if ((rc = reinit_begin(h, rs, s))) return rc;
        for (int it = 0; it < iterations; ++it) {
            VLS_HIP(launch_cubic_reinit_stage(t0, F, nullptr, as.flags, as.lv, g, 0, h->over, rs.counts, rs.seq++, s));
            VLS_HIP(launch_cubic_reinit_stage(t1, t0, F, as.flags, as.lv, g, 1, h->over, rs.counts, rs.seq++, s));
            VLS_HIP(launch_cubic_reinit_stage(F, t1, F, as.flags, as.lv, g, 2, h->over, rs.counts, rs.seq++, s));
        }
        return VLS_OK;
    });
}

// ENORK2Extend (+SD/@SDF3/ENORK2Extend.m:9-81) and WENORK3Extend
// (+SD/@SDF3/WENORK3Extend.m:9-98): distances of F (quadratic ENO or cubic)
// into the auxiliary distance set, v = sign(F) grad/max(|grad|,1e-6) from the
// variant's upwind normal, the interface values of C once, then Heun RK2 or
// Shu-Osher RK3 steps of C in place.  ENO-RK2 runs one point kernel per stage
// (full 6-row interface values in h->lcp); WENO-RK3 runs the fused stage
// kernels of the WENO5 path with cubic ENO at the listed points (list-order
// interface values and velocities in h->lcp, h->lvel).
static int variant_extend(vls_sdf3* h, const double* F, double* C, int iterations, int weno, hipStream_t s)
{
    VLS_REQUIRE(h && F && C, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    VLS_REQUIRE((const void*)F != (const void*)C, "C must not alias F");
    return with_graph(h->graphs, make_key(22 + weno, F, C, iterations), s, [&](hipStream_t cs) -> int {
        const Geom& g = h->g;
        const int64_t n = g.n();
        const DSet as = aux_set(h);
        int rc = set_distance_into<double>(h, weno ? 3 : 2, as, nullptr, F, nullptr, true, cs);
        if (rc) return rc;
        const Dist6c dc = constify(h->ad);
        double *vx = h->v[0], *vy = h->v[1], *vz = h->v[2];
        if (weno) VLS_HIP(launch_wenork3_upwind_normal(vx, vy, vz, F, dc, g, cs));
        else VLS_HIP(launch_enork2_upwind_normal(vx, vy, vz, F, dc, g, cs));
        VLS_HIP(launch_extension_velocity<double>(vx, vy, vz, F, 1e-6, g, cs));
        double *t0 = h->t[0], *t1 = h->t[1];
        if (weno) {
            VLS_HIP(launch_list_gather_vel<double>(h->lvel, as.lv, vx, vy, vz, g, cs));
            VLS_HIP(launch_list_interp_cubic(h->lcp, as.lv, C, g, cs));
            for (int it = 0; it < iterations; ++it) {
                VLS_HIP(launch_cubic_extend_stage(t0, C, nullptr, vx, vy, vz, as.lv, h->lcp, h->lvel, g, 0,
                                                  cs));
                VLS_HIP(launch_cubic_extend_stage(t1, t0, C, vx, vy, vz, as.lv, h->lcp, h->lvel, g, 1, cs));
                VLS_HIP(launch_cubic_extend_stage(C, t1, C, vx, vy, vz, as.lv, h->lcp, h->lvel, g, 2, cs));
            }
            return VLS_OK;
        }
        double* q = h->lcp;
        const Dist6 cp = mk6(q, q + n, q + 2 * n, q + 3 * n, q + 4 * n, q + 5 * n);
        VLS_HIP(launch_enork2_boundary_interpolate(cp, dc, C, g, cs));
        const Dist6c cpc = constify(cp);
        for (int it = 0; it < iterations; ++it) {
            VLS_HIP(launch_enork2_extend_stage(t0, C, nullptr, vx, vy, vz, h->aflags, dc, cpc, g, 0, cs));
            VLS_HIP(launch_enork2_extend_stage(C, t0, C, vx, vy, vz, h->aflags, dc, cpc, g, 1, cs));
        }
        return VLS_OK;
    });
}

int vls_sdf3_enork2_extend(vls_sdf3_t h, const double* F, double* C, int iterations, vls_stream_t stream)
{
    return variant_extend(h, F, C, iterations, 0, S(stream));
}

int vls_sdf3_wenork3_extend(vls_sdf3_t h, const double* F, double* C, int iterations, vls_stream_t stream)
{
    return variant_extend(h, F, C, iterations, 1, S(stream));
}

int vls_sdf3_weno5_rk3_extend(vls_sdf3_t h, const double* F, double* C, int iterations, vls_stream_t stream)
{
    return sdf3_weno5_rk3_extend<double>(h, F, C, iterations, S(stream));
}

int vls_sdf3_weno5_rk3_extend_f32(vls_sdf3_t h, const float* F, float* C, int iterations, vls_stream_t stream)
{
    return sdf3_weno5_rk3_extend<float>(h, F, C, iterations, S(stream));
}

int vls_sdf3_surface_redistance(vls_sdf3_t h, const double* F, double* A, int iter1, int iter2, vls_stream_t stream)
{
    return sdf3_surface_redistance<double>(h, F, A, iter1, iter2, S(stream));
}

int vls_sdf3_surface_redistance_f32(vls_sdf3_t h, const float* F, float* A, int iter1, int iter2,
                                    vls_stream_t stream)
{
    return sdf3_surface_redistance<float>(h, F, A, iter1, iter2, S(stream));
}

int vls_sdf3_advect_rk3(vls_sdf3_t h, double* F, const double* vx, const double* vy, const double* vz, double dt,
                        int steps, vls_stream_t stream)
{
    return sdf3_advect_rk3<double>(h, F, vx, vy, vz, dt, steps, S(stream));
}

int vls_sdf3_advect_rk3_f32(vls_sdf3_t h, float* F, const float* vx, const float* vy, const float* vz, double dt,
                            int steps, vls_stream_t stream)
{
    return sdf3_advect_rk3<float>(h, F, vx, vy, vz, dt, steps, S(stream));
}

int vls_sdf3_max_speed(vls_sdf3_t h, const double* vx, const double* vy, const double* vz, double* out,
                       vls_stream_t stream)
{
    return sdf3_max_speed<double>(h, vx, vy, vz, out, S(stream));
}

int vls_sdf3_max_speed_f32(vls_sdf3_t h, const float* vx, const float* vy, const float* vz, double* out,
                           vls_stream_t stream)
{
    return sdf3_max_speed<float>(h, vx, vy, vz, out, S(stream));
}

int vls_sdf3_area_volume_towers(vls_sdf3_t h, const double* F, double* area, double* volume, vls_stream_t stream)
{
    VLS_REQUIRE(area && volume, "null argument");
    return area_volume<double>(h, 0, F, 0.0, nullptr, area, volume, S(stream));
}

int vls_sdf3_tension_pressure(vls_sdf3_t h, const double* F, const double* H, const double* Vb, const double* dirac,
                              const double* gradmag, const double* heaviside, double area0, double volume0,
                              double dt, double cfl, double* speed, double* out10, vls_stream_t stream)
{
    VLS_REQUIRE(h, "null handle");
    VLS_REQUIRE(!any_null(F, H, Vb, dirac, gradmag, heaviside, speed, out10), "null pointer");
    VLS_REQUIRE(std::isfinite(area0) && std::isfinite(volume0) && std::isfinite(dt) && std::isfinite(cfl),
                "area0, volume0, dt and cfl must be finite");
    VLS_REQUIRE(dt > 0 || cfl > 0, "give a time step dt > 0 or a CFL number cfl > 0");
    const double ds = (h->g.dx + h->g.dy + h->g.dz) / 3;
    VLS_HIP(launch_tension_pressure(out10, h->partial, speed, F, H, Vb, dirac, gradmag, heaviside, h->g, ds, area0,
                                    volume0, dt, cfl, S(stream)));
    return VLS_OK;
}

int vls_sdf3_area_volume_cosine(vls_sdf3_t h, const double* F, double eps, double* area, double* volume,
                                vls_stream_t stream)
{
    VLS_REQUIRE(area && volume, "null argument");
    return area_volume<double>(h, 1, F, eps, nullptr, area, volume, S(stream));
}

int vls_sdf3_area_volume_device(vls_sdf3_t h, const double* F, int kind, double eps, double* out2,
                                vls_stream_t stream)
{
    VLS_REQUIRE(out2, "null argument");
    return area_volume<double>(h, kind, F, eps, out2, nullptr, nullptr, S(stream));
}

int vls_sdf3_area_volume_device_f32(vls_sdf3_t h, const float* F, int kind, double eps, double* out2,
                                    vls_stream_t stream)
{
    VLS_REQUIRE(out2, "null argument");
    return area_volume<float>(h, kind, F, eps, out2, nullptr, nullptr, S(stream));
}

int vls_sdf3_normal_velocity(vls_sdf3_t h, const double* F, const double* vx, const double* vy, const double* vz,
                             double* out, vls_stream_t stream)
{
    return sdf3_normal_velocity<double>(h, F, vx, vy, vz, out, S(stream));
}

int vls_sdf3_normal_velocity_f32(vls_sdf3_t h, const float* F, const float* vx, const float* vy, const float* vz,
                                 float* out, vls_stream_t stream)
{
    return sdf3_normal_velocity<float>(h, F, vx, vy, vz, out, S(stream));
}

}  // extern "C"

// ---------------------------------------------------------------------------
// 4. batches of grids
// ---------------------------------------------------------------------------
struct vls_batch {
    int ngrids = 0, ntiles = 0;
    int64_t total = 0;
    int blocks = 0;  // persistent blocks per tile-loop launch
    int tr = 64;     // rows per tile
    int zd = 8;      // pages per work item of the z-marching bulk stage kernels
    int nitems = 0;
    ZItem* d_items = nullptr;
    std::vector<BGrid> grids;
    std::vector<int64_t> offs;
    BGrid* d_grids = nullptr;
    int* d_tile_grid = nullptr;
    int64_t* d_offs = nullptr;
    double* pool = nullptr;  // list data and work fields
    uint8_t* flags = nullptr;
    int* int_pool = nullptr;
    ListView lv{};
    int* over = nullptr;
    int* list_grid = nullptr;  // grid of each interface-list entry
    int* counts = nullptr;     // [0] list count, [3] candidate count of the list build, [4..5] deferred counts
    double* lcp = nullptr;
    double* lvel = nullptr;
    double* t[2] = {nullptr, nullptr};
    double* v[3] = {nullptr, nullptr, nullptr};
    double* partial = nullptr;  // per-tile partial sums (4 per tile of either shape)
    GraphCache graphs;
    bool list_ready = false;  // a set_distance (or step) has built the interface list

    BatchView view() const
    {
        return BatchView{d_grids, d_tile_grid, d_offs, ngrids, ntiles, blocks, tr, d_items, nitems, list_grid};
    }
    ~vls_batch()
    {
        (void)hipFree(d_items);
        (void)hipFree(d_grids);
        (void)hipFree(d_tile_grid);
        (void)hipFree(d_offs);
        (void)hipFree(pool);
        (void)hipFree(flags);
        (void)hipFree(int_pool);
        (void)hipFree(partial);
    }
};

namespace {

// persistent blocks of a tile-loop launch: 8 blocks of 256 threads per CU
// (the occupancy of the bulk kernels) on every CU, a multiple of 8 so that
// each XCD gets the same share, never more than there are tiles (rounded up
// to a multiple of 8)
int default_blocks(int ntiles)
{
    int b = 8 * cu_count();
    const int cap = (ntiles + 7) / 8 * 8;
    if (b > cap) b = cap;
    return b < 8 ? 8 : b;
}

// the tile list of shape `tr` (host descriptors, device descriptors and tile->grid map)
int build_tiles(vls_batch* b, int tr)
{
    std::vector<int> tile_grid;
    int64_t tiles = 0;
    for (int g = 0; g < b->ngrids; ++g) {
        BGrid& G = b->grids[(size_t)g];
        G.ntiles = grid_tiles(G.rows, G.cols, G.pges, tr);
        G.tile0 = (int)tiles;
        tiles += G.ntiles;
        if (tiles >= ((int64_t)1 << 31)) return fail(VLS_ERR_INVALID, "too many tiles");
        tile_grid.insert(tile_grid.end(), (size_t)G.ntiles, g);
    }
    // z-march work list: per grid, per chunk of zd pages, the tile columns
    // (neighbouring columns adjacent, so consecutive blocks share halos)
    std::vector<ZItem> items;
    const int tc = 256 / tr;
    for (int g = 0; g < b->ngrids; ++g) {
        const BGrid& G = b->grids[(size_t)g];
        const int ntr = (G.rows + tr - 1) / tr, ntc = (G.cols + tc - 1) / tc;
        for (int p0 = 0; p0 < G.pges; p0 += b->zd)
            for (int cb = 0; cb < ntc; ++cb)
                for (int rb = 0; rb < ntr; ++rb)
                    items.push_back(ZItem{g, rb * tr, cb * tc, p0, std::min(p0 + b->zd, G.pges)});
        if (items.size() >= ((size_t)1 << 31)) return fail(VLS_ERR_INVALID, "too many work items");
    }
    (void)hipFree(b->d_tile_grid);
    b->d_tile_grid = nullptr;
    (void)hipFree(b->d_items);
    b->d_items = nullptr;
    if (hipMalloc((void**)&b->d_tile_grid, sizeof(int) * (size_t)tiles) != hipSuccess)
        return fail(VLS_ERR_ALLOC, "hipMalloc of the tile list failed");
    if (hipMalloc((void**)&b->d_items, sizeof(ZItem) * items.size()) != hipSuccess)
        return fail(VLS_ERR_ALLOC, "hipMalloc of the work list failed");
    VLS_HIP(hipMemcpy(b->d_tile_grid, tile_grid.data(), sizeof(int) * (size_t)tiles, hipMemcpyHostToDevice));
    VLS_HIP(hipMemcpy(b->d_items, items.data(), sizeof(ZItem) * items.size(), hipMemcpyHostToDevice));
    VLS_HIP(hipMemcpy(b->d_grids, b->grids.data(), sizeof(BGrid) * (size_t)b->ngrids, hipMemcpyHostToDevice));
    b->ntiles = (int)tiles;
    b->nitems = (int)items.size();
    b->tr = tr;
    return VLS_OK;
}

template <typename T>
T* btmp(vls_batch* b, int k)
{
    return reinterpret_cast<T*>(b->t[k]);
}
template <typename T>
T* bvel(vls_batch* b, int a)
{
    return reinterpret_cast<T*>(b->v[a]);
}

template <typename T>
int batch_set_distance(vls_batch* b, const T* F, const T* mask_src, hipStream_t s)
{
    VLS_HIP(launch_zero_ints(b->counts, 4, s));  // the list count and the candidate count (counts[3])
    VLS_HIP(launch_batch_build_list<T>(b->view(), F, mask_src, b->flags, b->lv, 6, b->over, b->counts + 3, s));
    return VLS_OK;
}

template <typename T>
int batch_reinit_loop(vls_batch* b, T* F, int iterations, hipStream_t s)
{
    const BatchView B = b->view();
    int* oc = b->counts + 4;
    T* t0 = btmp<T>(b, 0);
    T* t1 = btmp<T>(b, 1);
    VLS_HIP(launch_batch_set_mask_bit<T>(b->flags, F, b->total, s));
    VLS_HIP(launch_zero_ints(oc, 2, s));
    int seq = 0;
    for (int it = 0; it < iterations; ++it) {
        VLS_HIP(launch_batch_reinit_stage<T>(B, t0, F, nullptr, b->flags, b->lv, 0, b->over, oc, seq++, s));
        VLS_HIP(launch_batch_reinit_stage<T>(B, t1, t0, F, b->flags, b->lv, 1, b->over, oc, seq++, s));
        VLS_HIP(launch_batch_reinit_stage<T>(B, F, t1, F, b->flags, b->lv, 2, b->over, oc, seq++, s));
    }
    return VLS_OK;
}

template <typename T>
int batch_extend(vls_batch* b, const T* F, T* C, int iterations, hipStream_t s)
{
    const BatchView B = b->view();
    T *vx = bvel<T>(b, 0), *vy = bvel<T>(b, 1), *vz = bvel<T>(b, 2);
    T* t0 = btmp<T>(b, 0);
    T* t1 = btmp<T>(b, 1);
    VLS_HIP(launch_batch_extension_velocity<T>(B, vx, vy, vz, F, b->flags, b->lv, 1e-6, s));
    VLS_HIP(launch_batch_list_gather_vel<T>(b->lvel, b->lv, vx, vy, vz, s));
    VLS_HIP(launch_batch_list_interp<T>(B, b->lcp, b->lv, C, s));
    for (int it = 0; it < iterations; ++it) {
        VLS_HIP(launch_batch_extend_stage<T>(B, t0, C, nullptr, vx, vy, vz, b->lv, b->lcp, b->lvel, 0, s));
        VLS_HIP(launch_batch_extend_stage<T>(B, t1, t0, C, vx, vy, vz, b->lv, b->lcp, b->lvel, 1, s));
        VLS_HIP(launch_batch_extend_stage<T>(B, C, t1, C, vx, vy, vz, b->lv, b->lcp, b->lvel, 2, s));
    }
    return VLS_OK;
}

template <typename T>
int batch_advect(vls_batch* b, T* F, const T* vx, const T* vy, const T* vz, const double* dt, int steps,
                 hipStream_t s)
{
    const BatchView B = b->view();
    T* t0 = btmp<T>(b, 0);
    T* t1 = btmp<T>(b, 1);
    for (int it = 0; it < steps; ++it) {
        VLS_HIP(launch_batch_advect_stage<T>(B, t0, F, nullptr, vx, vy, vz, dt, 0, s));
        VLS_HIP(launch_batch_advect_stage<T>(B, t1, t0, F, vx, vy, vz, dt, 1, s));
        VLS_HIP(launch_batch_advect_stage<T>(B, F, t1, F, vx, vy, vz, dt, 2, s));
    }
    return VLS_OK;
}

// one vesicle step of every grid: TVD-RK3 advection of F, setDistance, the
// normal velocity u.N of F into Vn and its WENO5-RK3 extension, a WENO5-RK3
// reinitialization of F when reinit_iters > 0, and the cosine area/volume of
// every grid into av (2 per grid)
template <typename T>
int batch_step(vls_batch* b, T* F, T* Vn, const T* vx, const T* vy, const T* vz, const double* dt, int ext_iters,
               int reinit_iters, double eps_factor, double* av, hipStream_t s)
{
    int rc;
    if ((rc = batch_advect<T>(b, F, vx, vy, vz, dt, 1, s))) return rc;
    if ((rc = batch_set_distance<T>(b, F, (const T*)nullptr, s))) return rc;
    VLS_HIP(launch_batch_normal_velocity<T>(b->view(), Vn, F, vx, vy, vz, s));
    if ((rc = batch_extend<T>(b, F, Vn, ext_iters, s))) return rc;
    if (reinit_iters > 0 && (rc = batch_reinit_loop<T>(b, F, reinit_iters, s))) return rc;
    VLS_HIP(launch_batch_area_volume_cosine<T>(b->view(), av, b->partial, F, eps_factor, s));
    return VLS_OK;
}

// WENORK3Reinitialization of every grid (+SD/@SDF3/WENORK3Reinitialization.m:3-61):
// cubic distances of F and mask F < 0 (:17-26), then Shu-Osher RK3 (:30-57)
// on the fused stage launches with the cubic ENO at the listed and deferred
// points, as vls_sdf3_wenork3_reinit does for one grid.  Leaves the list
// holding the cubic distances.
int batch_wenork3_reinit(vls_batch* b, double* F, int iterations, hipStream_t s)
{
    const BatchView B = b->view();
    int* oc = b->counts + 4;
    double* t0 = btmp<double>(b, 0);
    double* t1 = btmp<double>(b, 1);
    VLS_HIP(launch_zero_ints(b->counts, 4, s));
    VLS_HIP(launch_batch_build_list<double>(B, F, F, b->flags, b->lv, 3, b->over, b->counts + 3, s));
    VLS_HIP(launch_zero_ints(oc, 2, s));
    int seq = 0;
    for (int it = 0; it < iterations; ++it) {
        VLS_HIP(launch_batch_reinit_stage<double>(B, t0, F, nullptr, b->flags, b->lv, 0, b->over, oc, seq++, s,
                                                  IFACE_CUBIC));
        VLS_HIP(launch_batch_reinit_stage<double>(B, t1, t0, F, b->flags, b->lv, 1, b->over, oc, seq++, s,
                                                  IFACE_CUBIC));
        VLS_HIP(launch_batch_reinit_stage<double>(B, F, t1, F, b->flags, b->lv, 2, b->over, oc, seq++, s,
                                                  IFACE_CUBIC));
    }
    return VLS_OK;
}

// WENORK3Extend (weno = 1, +SD/@SDF3/WENORK3Extend.m:9-98) and ENORK2Extend
// (weno = 0, ENORK2Extend.m:9-81) of C off F on every grid, as
// variant_extend does for one grid: cubic or quadratic-ENO distances of F,
// the variant's upwind normal made the extension velocity, the interface
// values of C once, then Shu-Osher RK3 on the fused stage launches (cubic
// ENO at the listed points) or Heun RK2 stages.  Leaves the list holding the
// variant's distances.
int batch_variant_extend(vls_batch* b, const double* F, double* C, int iterations, int weno, hipStream_t s)
{
    const BatchView B = b->view();
    double *vx = bvel<double>(b, 0), *vy = bvel<double>(b, 1), *vz = bvel<double>(b, 2);
    double* t0 = btmp<double>(b, 0);
    double* t1 = btmp<double>(b, 1);
    const int scheme = weno ? 1 : 2;  // the cubic ENO, the quadratic ENO
    VLS_HIP(launch_zero_ints(b->counts, 4, s));
    VLS_HIP(launch_batch_build_list<double>(B, F, nullptr, b->flags, b->lv, weno ? 3 : 2, b->over, b->counts + 3,
                                            s));
    VLS_HIP(launch_batch_extension_velocity<double>(B, vx, vy, vz, F, b->flags, b->lv, 1e-6, s, scheme));
    VLS_HIP(launch_batch_list_interp<double>(B, b->lcp, b->lv, C, s, scheme));
    if (weno) {
        VLS_HIP(launch_batch_list_gather_vel<double>(b->lvel, b->lv, vx, vy, vz, s));
        for (int it = 0; it < iterations; ++it) {
            VLS_HIP(launch_batch_extend_stage<double>(B, t0, C, nullptr, vx, vy, vz, b->lv, b->lcp,
                                                      b->lvel, 0, s, IFACE_CUBIC));
            VLS_HIP(launch_batch_extend_stage<double>(B, t1, t0, C, vx, vy, vz, b->lv, b->lcp, b->lvel, 1,
                                                      s, IFACE_CUBIC));
            VLS_HIP(launch_batch_extend_stage<double>(B, C, t1, C, vx, vy, vz, b->lv, b->lcp, b->lvel, 2,
                                                      s, IFACE_CUBIC));
        }
        return VLS_OK;
    }
    for (int it = 0; it < iterations; ++it) {
        VLS_HIP(launch_batch_eno2_extend_stage(B, t0, C, nullptr, vx, vy, vz, b->flags, b->lv, b->lcp, 0, s));
        VLS_HIP(launch_batch_eno2_extend_stage(B, C, t0, C, vx, vy, vz, b->flags, b->lv, b->lcp, 1, s));
    }
    return VLS_OK;
}

}  // namespace

extern "C" {

int vls_batch_create(vls_batch_t* out, int ngrids, const int* rows, const int* cols, const int* pges,
                     const double* dx, const double* dy, const double* dz)
{
    VLS_REQUIRE(out != nullptr, "null handle pointer");
    *out = nullptr;
    VLS_REQUIRE(ngrids > 0, "a batch needs at least one grid");
    VLS_REQUIRE(!any_null(rows, cols, pges, dx, dy, dz), "null argument");
    vls_batch* b = new (std::nothrow) vls_batch();
    if (!b) return fail(VLS_ERR_ALLOC, "host allocation failed");
    b->ngrids = ngrids;
    b->offs.resize((size_t)ngrids + 1);
    int64_t off = 0, max_tiles = 0, tiles64 = 0, tiles32 = 0;
    for (int g = 0; g < ngrids; ++g) {
        int rc = check_geom(rows[g], cols[g], pges[g], dx[g], dy[g], dz[g]);
        if (rc) {
            delete b;
            return rc;
        }
        BGrid G{};
        G.rows = rows[g];
        G.cols = cols[g];
        G.pges = pges[g];
        G.off = off;
        G.dx = dx[g];
        G.dy = dy[g];
        G.dz = dz[g];
        G.inv[0] = 1.0 / dx[g];
        G.inv[1] = 1.0 / dy[g];
        G.inv[2] = 1.0 / dz[g];
        const double m = std::min(std::min(dx[g], dx[g]), std::min(std::min(dy[g], dy[g]), std::min(dz[g], dz[g])));
        G.dt_bulk = 0.3 * m;
        set_jp(G.jd, G.jf, dx[g], dy[g], dz[g]);
        b->grids.push_back(G);
        b->offs[(size_t)g] = off;
        off += (int64_t)rows[g] * cols[g] * pges[g];
        tiles64 += grid_tiles(rows[g], cols[g], pges[g], 64);
        tiles32 += grid_tiles(rows[g], cols[g], pges[g], 32);
        if (off >= ((int64_t)1 << 31)) {
            delete b;
            return fail(VLS_ERR_INVALID, "a batch holds fewer than 2^31 points");
        }
    }
    max_tiles = std::max(tiles64, tiles32);
    b->offs[(size_t)ngrids] = off;
    b->total = off;
    const int64_t n = off;
    // list rows (6 dist + 1 dt + 6 cp + 3 vel) + work fields (2 t + 3 v), one allocation
    const size_t dbl = (size_t)n * (6 + 1 + 6 + 3 + 2 + 3);
    if (hipMalloc((void**)&b->pool, sizeof(double) * dbl) != hipSuccess ||
        hipMalloc((void**)&b->partial, sizeof(double) * ((size_t)max_tiles * 4 + 8)) != hipSuccess ||
        hipMalloc((void**)&b->flags, (size_t)n) != hipSuccess ||
        hipMalloc((void**)&b->int_pool, sizeof(int) * ((size_t)n * 3 + 16)) != hipSuccess ||
        hipMalloc((void**)&b->d_grids, sizeof(BGrid) * (size_t)ngrids) != hipSuccess ||
        hipMalloc((void**)&b->d_offs, sizeof(int64_t) * ((size_t)ngrids + 1)) != hipSuccess) {
        delete b;
        return fail(VLS_ERR_ALLOC, "hipMalloc of the batch (%lld points) failed", (long long)n);
    }
    if (hipMemcpy(b->d_offs, b->offs.data(), sizeof(int64_t) * ((size_t)ngrids + 1), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemset(b->int_pool, 0, sizeof(int) * ((size_t)n * 3 + 16)) != hipSuccess ||
        hipMemset(b->flags, 0, (size_t)n) != hipSuccess) {
        delete b;
        return fail(VLS_ERR_HIP, "copying the batch descriptors failed");
    }
    // rows of 32 fill the 64-row tiles of every grid only half: shape 32 when
    // that needs fewer tiles than 64 (the autotuner may change it)
    int rc = build_tiles(b, tiles32 < tiles64 ? 32 : 64);
    if (rc) {
        delete b;
        return rc;
    }
    b->blocks = default_blocks(b->ntiles);
    double* p = b->pool;
    b->lv.dist = p; p += 6 * n;
    b->lv.dt = p; p += n;
    b->lcp = p; p += 6 * n;
    b->lvel = p; p += 3 * n;
    for (int k = 0; k < 2; ++k) { b->t[k] = p; p += n; }
    for (int a = 0; a < 3; ++a) { b->v[a] = p; p += n; }
    b->lv.idx = b->int_pool;
    b->over = b->int_pool + n;
    b->list_grid = b->int_pool + 2 * n;
    b->counts = b->int_pool + 3 * n;
    b->lv.count = b->counts;
    b->lv.cap = (int)n;
    *out = b;
    return VLS_OK;
}

int vls_batch_destroy(vls_batch_t b)
{
    delete b;
    return VLS_OK;
}

int vls_batch_layout(vls_batch_t b, int64_t* offsets, int* ntiles)
{
    VLS_REQUIRE(b && offsets, "null argument");
    for (size_t g = 0; g < b->offs.size(); ++g) offsets[g] = b->offs[g];
    if (ntiles) *ntiles = b->ntiles;
    return VLS_OK;
}

int vls_batch_set_blocks(vls_batch_t b, int blocks)
{
    VLS_REQUIRE(b, "null handle");
    VLS_REQUIRE(blocks >= 0 && blocks % 8 == 0, "blocks must be a non-negative multiple of 8 (0: default)");
    b->blocks = blocks ? blocks : default_blocks(b->ntiles);
    b->graphs.clear();  // the graphs hold the old launch shape
    return VLS_OK;
}

int vls_batch_get_blocks(vls_batch_t b, int* blocks)
{
    VLS_REQUIRE(b && blocks, "null argument");
    *blocks = b->blocks;
    return VLS_OK;
}

int vls_batch_set_tile_rows(vls_batch_t b, int tile_rows)
{
    VLS_REQUIRE(b, "null handle");
    VLS_REQUIRE(tile_rows == 32 || tile_rows == 64, "tile rows must be 32 or 64");
    if (tile_rows == b->tr) return VLS_OK;
    VLS_HIP(hipDeviceSynchronize());  // no launch of this batch may still read the old tile list
    const int rc = build_tiles(b, tile_rows);
    if (rc) return rc;
    b->blocks = std::min(b->blocks, (b->ntiles + 7) / 8 * 8);
    if (b->blocks < 8) b->blocks = 8;
    b->graphs.clear();
    return VLS_OK;
}

int vls_batch_get_tile_rows(vls_batch_t b, int* tile_rows)
{
    VLS_REQUIRE(b && tile_rows, "null argument");
    *tile_rows = b->tr;
    return VLS_OK;
}

int vls_batch_set_zmarch_depth(vls_batch_t b, int depth)
{
    VLS_REQUIRE(b, "null handle");
    VLS_REQUIRE(depth > 0, "depth must be positive");
    if (depth == b->zd) return VLS_OK;
    VLS_HIP(hipDeviceSynchronize());  // no launch of this batch may still read the old work list
    b->zd = depth;
    const int rc = build_tiles(b, b->tr);
    if (rc) return rc;
    b->graphs.clear();
    return VLS_OK;
}

int vls_batch_get_zmarch_depth(vls_batch_t b, int* depth)
{
    VLS_REQUIRE(b && depth, "null argument");
    *depth = b->zd;
    return VLS_OK;
}

int vls_batch_wenork3_reinit(vls_batch_t b, double* F, int iterations, vls_stream_t stream)
{
    VLS_REQUIRE(b && F, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    const int rc = with_graph(b->graphs, make_key(14, F, iterations), S(stream),
                              [&](hipStream_t cs) -> int { return batch_wenork3_reinit(b, F, iterations, cs); });
    b->list_ready = false;  // the list holds the cubic distances now
    return rc;
}

int vls_batch_wenork3_extend(vls_batch_t b, const double* F, double* C, int iterations, vls_stream_t stream)
{
    VLS_REQUIRE(b && F && C, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    VLS_REQUIRE((const void*)F != (const void*)C, "C must not alias F");
    const int rc = with_graph(b->graphs, make_key(15, F, C, iterations), S(stream),
                              [&](hipStream_t cs) -> int { return batch_variant_extend(b, F, C, iterations, 1, cs); });
    b->list_ready = false;
    return rc;
}

int vls_batch_enork2_extend(vls_batch_t b, const double* F, double* C, int iterations, vls_stream_t stream)
{
    VLS_REQUIRE(b && F && C, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    VLS_REQUIRE((const void*)F != (const void*)C, "C must not alias F");
    const int rc = with_graph(b->graphs, make_key(16, F, C, iterations), S(stream),
                              [&](hipStream_t cs) -> int { return batch_variant_extend(b, F, C, iterations, 0, cs); });
    b->list_ready = false;
    return rc;
}

#define VLS_BATCH_ENTRIES(T, SFX)                                                                                   \
    int vls_batch_set_distance##SFX(vls_batch_t b, const T* F, vls_stream_t stream)                                \
    {                                                                                                              \
        VLS_REQUIRE(b && F, "null argument");                                                                      \
        const int rc = batch_set_distance<T>(b, F, (const T*)nullptr, S(stream));                                  \
        if (rc == VLS_OK) b->list_ready = true;                                                                    \
        return rc;                                                                                                 \
    }                                                                                                              \
    int vls_batch_weno5_rk3_reinit##SFX(vls_batch_t b, T* F, int iterations, vls_stream_t stream)                  \
    {                                                                                                              \
        VLS_REQUIRE(b && F, "null argument");                                                                      \
        VLS_REQUIRE(iterations >= 0, "negative iteration count");                                                  \
        VLS_REQUIRE(b->list_ready, "set_distance must run before the reinitialization");                           \
        return with_graph(b->graphs, make_key(11, sizeof(T), F, iterations), S(stream),                            \
                          [&](hipStream_t cs) -> int { return batch_reinit_loop<T>(b, F, iterations, cs); });      \
    }                                                                                                              \
    int vls_batch_weno5_rk3_extend##SFX(vls_batch_t b, const T* F, T* C, int iterations, vls_stream_t stream)      \
    {                                                                                                              \
        VLS_REQUIRE(b && F && C, "null argument");                                                                 \
        VLS_REQUIRE(iterations >= 0, "negative iteration count");                                                  \
        VLS_REQUIRE((const void*)F != (const void*)C, "C must not alias F");                                       \
        VLS_REQUIRE(b->list_ready, "set_distance must run before the extension");                                  \
        return with_graph(b->graphs, make_key(12, sizeof(T), F, C, iterations), S(stream),                         \
                          [&](hipStream_t cs) -> int { return batch_extend<T>(b, F, C, iterations, cs); });        \
    }                                                                                                              \
    int vls_batch_advect_rk3##SFX(vls_batch_t b, T* F, const T* vx, const T* vy, const T* vz, const double* dt,    \
                                  int steps, vls_stream_t stream)                                                  \
    {                                                                                                              \
        VLS_REQUIRE(b && F && vx && vy && vz && dt, "null argument");                                              \
        VLS_REQUIRE(steps >= 0, "negative step count");                                                            \
        return with_graph(b->graphs, make_key(13, sizeof(T), F, vx, vy, vz, dt, steps), S(stream),                 \
                          [&](hipStream_t cs) -> int { return batch_advect<T>(b, F, vx, vy, vz, dt, steps, cs); });\
    }                                                                                                              \
    int vls_batch_normal_velocity##SFX(vls_batch_t b, const T* F, const T* vx, const T* vy, const T* vz, T* out,   \
                                       vls_stream_t stream)                                                        \
    {                                                                                                              \
        VLS_REQUIRE(b && F && vx && vy && vz && out, "null argument");                                             \
        VLS_REQUIRE((const void*)out != (const void*)F, "out must not alias F");                                   \
        VLS_HIP(launch_batch_normal_velocity<T>(b->view(), out, F, vx, vy, vz, S(stream)));                       \
        return VLS_OK;                                                                                             \
    }                                                                                                              \
    int vls_batch_area_volume_cosine##SFX(vls_batch_t b, const T* F, double eps_factor, double* out2n,             \
                                          vls_stream_t stream)                                                     \
    {                                                                                                              \
        VLS_REQUIRE(b && F && out2n, "null argument");                                                             \
        VLS_REQUIRE(eps_factor > 0, "eps_factor must be positive");                                                \
        VLS_HIP(launch_batch_area_volume_cosine<T>(b->view(), out2n, b->partial, F, eps_factor, S(stream)));      \
        return VLS_OK;                                                                                             \
    }                                                                                                              \
    int vls_batch_cfl_dt##SFX(vls_batch_t b, const T* vx, const T* vy, const T* vz, double cfl, double* dt,        \
                              vls_stream_t stream)                                                                 \
    {                                                                                                              \
        VLS_REQUIRE(b && vx && vy && vz && dt, "null argument");                                                   \
        VLS_REQUIRE(cfl > 0, "cfl must be positive");                                                             \
        VLS_HIP(launch_batch_cfl_dt<T>(b->view(), dt, b->partial, vx, vy, vz, cfl, S(stream)));                   \
        return VLS_OK;                                                                                             \
    }                                                                                                              \
    int vls_batch_curl_velocity##SFX(vls_batch_t b, const double* modes, int nmodes, const double* lo, T* vx,      \
                                     T* vy, T* vz, vls_stream_t stream)                                            \
    {                                                                                                              \
        VLS_REQUIRE(b && modes && lo && vx && vy && vz, "null argument");                                          \
        VLS_REQUIRE(nmodes >= 0, "negative mode count");                                                           \
        VLS_HIP(launch_batch_curl_velocity<T>(b->view(), vx, vy, vz, modes, nmodes, lo, S(stream)));              \
        return VLS_OK;                                                                                             \
    }                                                                                                              \
    int vls_batch_ellipsoid##SFX(vls_batch_t b, const double* axes, const double* lo, T* F, vls_stream_t stream)   \
    {                                                                                                              \
        VLS_REQUIRE(b && axes && lo && F, "null argument");                                                        \
        VLS_HIP(launch_batch_ellipsoid<T>(b->view(), F, axes, lo, S(stream)));                                    \
        return VLS_OK;                                                                                             \
    }                                                                                                              \
    int vls_batch_step##SFX(vls_batch_t b, T* F, T* Vn, const T* vx, const T* vy, const T* vz, const double* dt,   \
                            int ext_iters, int reinit_iters, double eps_factor, double* av, vls_stream_t stream)   \
    {                                                                                                              \
        VLS_REQUIRE(b && F && Vn && vx && vy && vz && dt && av, "null argument");                                  \
        VLS_REQUIRE(ext_iters >= 0 && reinit_iters >= 0, "negative iteration count");                              \
        VLS_REQUIRE(eps_factor > 0, "eps_factor must be positive");                                                \
        VLS_REQUIRE((const void*)F != (const void*)Vn, "Vn must not alias F");                                     \
        const int rc = with_graph(                                                                                 \
            b->graphs, make_key(14, sizeof(T), F, Vn, vx, vy, vz, dt, ext_iters, reinit_iters, eps_factor, av),    \
            S(stream), [&](hipStream_t cs) -> int {                                                                \
                return batch_step<T>(b, F, Vn, vx, vy, vz, dt, ext_iters, reinit_iters, eps_factor, av, cs);       \
            });                                                                                                    \
        if (rc == VLS_OK) b->list_ready = true;                                                                    \
        return rc;                                                                                                 \
    }

VLS_BATCH_ENTRIES(double, )
VLS_BATCH_ENTRIES(float, _f32)
#undef VLS_BATCH_ENTRIES

}  // extern "C"

// ---------------------------------------------------------------------------
// 5. autotuning of the launch shape
// ---------------------------------------------------------------------------
// Times `iterations` WENO5-RK3 extension iterations (the dominant operator)
// of a scratch copy of F, off F's own interface, and as many WENO5-RK3
// reinitialization iterations of another copy, for every candidate launch
// shape and keeps the shape with the smallest sum.  The results of every method are the same for
// every shape; only the time differs.  Candidates: tile rows 64 and 32
// (LDS tile 70 x 10 or 38 x 14 values) times the z-march depths (pages per
// block or work item): for one grid the default depth, half and twice it;
// for a batch 4, 8, 16 and 32 (the 512-vesicle batch of edges 32..96
// measured 0.4-0.7% faster at 32 than at 16 in FP64 and FP32).
namespace {

struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventTimer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

template <typename T>
int batch_autotune(vls_batch* b, const T* F, int iterations, hipStream_t s, int* tr_out, int* blocks_out)
{
    // candidates rebuild the work list the cached graphs point into: drop
    // them first, and on any failure go back to the shape the batch had
    b->graphs.clear();
    const int tr0 = b->tr, zd0 = b->zd, blocks0 = b->blocks;
    auto restore = [&](int rc) -> int {
        b->zd = zd0;
        if (build_tiles(b, tr0) == VLS_OK) b->blocks = blocks0;
        b->graphs.clear();
        return rc;
    };
    T* C = nullptr;
    if (hipMalloc((void**)&C, sizeof(T) * (size_t)b->total) != hipSuccess)
        return restore(fail(VLS_ERR_ALLOC, "hipMalloc of the autotuning field failed"));
    EventTimer ev;
    int rc = VLS_OK;
    float best = 1e30f;
    int best_tr = b->tr, best_zd = b->zd;
    const int trs[2] = {64, 32};
    const int depths[4] = {4, 8, 16, 32};
    if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess) rc = fail(VLS_ERR_HIP, "hipEventCreate");
    for (int i = 0; i < 2 && rc == VLS_OK; ++i) {
        for (int j = 0; j < 4 && rc == VLS_OK; ++j) {
            b->zd = depths[j];
            if ((rc = build_tiles(b, trs[i]))) break;
            b->blocks = default_blocks(b->ntiles);
            if (hipMemcpyAsync(C, F, sizeof(T) * (size_t)b->total, hipMemcpyDeviceToDevice, s) != hipSuccess) {
                rc = fail(VLS_ERR_HIP, "hipMemcpyAsync");
                break;
            }
            if ((rc = batch_set_distance<T>(b, F, (const T*)nullptr, s))) break;
            if ((rc = batch_extend<T>(b, F, C, 1, s))) break;
            if (hipEventRecord(ev.a, s) != hipSuccess) { rc = fail(VLS_ERR_HIP, "hipEventRecord"); break; }
            if ((rc = batch_extend<T>(b, F, C, iterations, s))) break;
            float ms = 0, msr = 0;
            if (hipEventRecord(ev.b, s) != hipSuccess || hipEventSynchronize(ev.b) != hipSuccess ||
                hipEventElapsedTime(&ms, ev.a, ev.b) != hipSuccess) {
                rc = fail(VLS_ERR_HIP, "timing the candidate failed");
                break;
            }
            // the reinitialization stages share the launch shape: time them on a copy of F
            if (hipMemcpyAsync(C, F, sizeof(T) * (size_t)b->total, hipMemcpyDeviceToDevice, s) != hipSuccess) {
                rc = fail(VLS_ERR_HIP, "hipMemcpyAsync");
                break;
            }
            if ((rc = batch_reinit_loop<T>(b, C, 1, s))) break;
            if (hipMemcpyAsync(C, F, sizeof(T) * (size_t)b->total, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                hipEventRecord(ev.a, s) != hipSuccess) {
                rc = fail(VLS_ERR_HIP, "hipMemcpyAsync");
                break;
            }
            if ((rc = batch_reinit_loop<T>(b, C, iterations, s))) break;
            if (hipEventRecord(ev.b, s) != hipSuccess || hipEventSynchronize(ev.b) != hipSuccess ||
                hipEventElapsedTime(&msr, ev.a, ev.b) != hipSuccess) {
                rc = fail(VLS_ERR_HIP, "timing the candidate failed");
                break;
            }
            ms += msr;
            if (ms < best) {
                best = ms;
                best_tr = trs[i];
                best_zd = depths[j];
            }
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(C);
    if (rc) return restore(rc);
    b->zd = best_zd;
    if ((rc = build_tiles(b, best_tr))) return restore(rc);
    b->blocks = default_blocks(b->ntiles);
    const int best_blocks = b->blocks;
    b->graphs.clear();
    b->list_ready = true;  // the last candidate ran set_distance on F
    if (tr_out) *tr_out = best_tr;
    if (blocks_out) *blocks_out = best_blocks;
    return VLS_OK;
}

template <typename T>
int sdf3_autotune_try(vls_sdf3* h, const T* F, int iterations, hipStream_t s, int* tr_out)
{
    EventTimer ev;
    VLS_HIP(hipEventCreate(&ev.a));
    VLS_HIP(hipEventCreate(&ev.b));
    T* C = tmp<T>(h, 2);
    const DSet ms = main_set(h);
    T *vx = vel<T>(h, 0), *vy = vel<T>(h, 1), *vz = vel<T>(h, 2);
    float best = 1e30f;
    int best_tr = h->g.tr, best_zd = h->g.zd;
    const int trs[2] = {64, 32};
    int rc;
    for (int i = 0; i < 2; ++i) {
        h->g.tr = trs[i];
        h->g.zd = 0;
        const int d0 = zmarch_depth(h->g);
        const int depths[3] = {d0, d0 > 1 ? d0 / 2 : 0, 2 * d0 <= h->g.pges ? 2 * d0 : 0};
        for (int j = 0; j < 3; ++j) {
            if (depths[j] <= 0) continue;
            h->g.zd = depths[j];
            const Geom& g = h->g;
            VLS_HIP(hipMemcpyAsync(C, F, sizeof(T) * (size_t)g.n(), hipMemcpyDeviceToDevice, s));
            if ((rc = set_distance_into<T>(h, 6, ms, h->min_dist, F, (const T*)nullptr, false, s))) return rc;
            h->dense_ready = false;
            VLS_HIP(launch_extension_normal<T>(vx, vy, vz, F, ms.flags, ms.lv, 1e-6, g, s));
            VLS_HIP(launch_list_gather_vel<T>(h->lvel, ms.lv, vx, vy, vz, g, s));
            if ((rc = extend_loop<T>(h, C, 1, ms, s))) return rc;
            VLS_HIP(hipEventRecord(ev.a, s));
            if ((rc = extend_loop<T>(h, C, iterations, ms, s))) return rc;
            VLS_HIP(hipEventRecord(ev.b, s));
            VLS_HIP(hipEventSynchronize(ev.b));
            float ms_ = 0, ms_r = 0;
            VLS_HIP(hipEventElapsedTime(&ms_, ev.a, ev.b));
            // the reinitialization stages share the launch shape: time them on a copy of F
            T* Fr = tmp<T>(h, 3);
            VLS_HIP(launch_set_mask_bit<T>(h->flags, F, g, s));
            VLS_HIP(hipMemcpyAsync(Fr, F, sizeof(T) * (size_t)g.n(), hipMemcpyDeviceToDevice, s));
            if ((rc = reinit_loop<T>(h, Fr, 1, ms, s))) return rc;
            VLS_HIP(hipMemcpyAsync(Fr, F, sizeof(T) * (size_t)g.n(), hipMemcpyDeviceToDevice, s));
            VLS_HIP(hipEventRecord(ev.a, s));
            if ((rc = reinit_loop<T>(h, Fr, iterations, ms, s))) return rc;
            VLS_HIP(hipEventRecord(ev.b, s));
            VLS_HIP(hipEventSynchronize(ev.b));
            VLS_HIP(hipEventElapsedTime(&ms_r, ev.a, ev.b));
            ms_ += ms_r;
            if (ms_ < best) {
                best = ms_;
                best_tr = trs[i];
                best_zd = depths[j];
            }
        }
    }
    h->g.tr = best_tr;
    h->g.zd = best_zd;
    h->graphs.clear();
    h->dist_ready = true;  // every candidate ran set_distance on F
    if (tr_out) *tr_out = best_tr;
    return VLS_OK;
}

// on a failure the handle keeps the shape it had (no cached graph survives)
template <typename T>
int sdf3_autotune(vls_sdf3* h, const T* F, int iterations, hipStream_t s, int* tr_out)
{
    const int tr0 = h->g.tr, zd0 = h->g.zd;
    h->graphs.clear();
    const int rc = sdf3_autotune_try<T>(h, F, iterations, s, tr_out);
    if (rc != VLS_OK) {
        h->g.tr = tr0;
        h->g.zd = zd0;
        h->graphs.clear();
    }
    return rc;
}

}  // namespace

extern "C" {

int vls_batch_autotune(vls_batch_t b, const double* F, int iterations, int* tile_rows, int* blocks,
                       vls_stream_t stream)
{
    VLS_REQUIRE(b && F, "null argument");
    VLS_REQUIRE(iterations > 0, "iterations must be positive");
    return batch_autotune<double>(b, F, iterations, S(stream), tile_rows, blocks);
}

int vls_batch_autotune_f32(vls_batch_t b, const float* F, int iterations, int* tile_rows, int* blocks,
                           vls_stream_t stream)
{
    VLS_REQUIRE(b && F, "null argument");
    VLS_REQUIRE(iterations > 0, "iterations must be positive");
    return batch_autotune<float>(b, F, iterations, S(stream), tile_rows, blocks);
}

int vls_sdf3_autotune(vls_sdf3_t h, const double* F, int iterations, int* tile_rows, vls_stream_t stream)
{
    VLS_REQUIRE(h && F, "null argument");
    VLS_REQUIRE(iterations > 0, "iterations must be positive");
    return sdf3_autotune<double>(h, F, iterations, S(stream), tile_rows);
}

int vls_sdf3_autotune_f32(vls_sdf3_t h, const float* F, int iterations, int* tile_rows, vls_stream_t stream)
{
    VLS_REQUIRE(h && F, "null argument");
    VLS_REQUIRE(iterations > 0, "iterations must be positive");
    return sdf3_autotune<float>(h, F, iterations, S(stream), tile_rows);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// 6. discrete adjoint and tangent (FP64, one grid)
// ---------------------------------------------------------------------------
namespace {

int ensure_adj(vls_sdf3* h, size_t len)
{
    if (h->adj_len >= len) return VLS_OK;
    VLS_HIP(hipDeviceSynchronize());  // the old workspace may still be in use
    (void)hipFree(h->adj);
    h->adj = nullptr;
    h->adj_len = 0;
    if (hipMalloc((void**)&h->adj, sizeof(double) * len) != hipSuccess)
        return fail(VLS_ERR_ALLOC, "hipMalloc of the adjoint workspace (%zu doubles) failed", len);
    h->adj_len = len;
    return VLS_OK;
}

int check_adj_box(const Geom& g)
{
    if (g.rows < 4 || g.cols < 4 || g.pges < 4) return fail(VLS_ERR_INVALID, "the adjoint needs at least 4 points per axis");
    return VLS_OK;
}

// stage coefficients of fn: 0, 3/4, 1/3
inline double stage_e(int stage) { return stage == 0 ? 0.0 : (stage == 1 ? 0.75 : 1.0 / 3.0); }

}  // namespace

extern "C" {

int vls_sdf3_adjoint_advect_rk3(vls_sdf3_t h, const double* F, const double* vx, const double* vy, const double* vz,
                                double dt, const double* lam_out, double* lam_in, double* gx, double* gy, double* gz,
                                vls_stream_t stream)
{
    VLS_REQUIRE(h && F && vx && vy && vz && lam_out && lam_in && gx && gy && gz, "null argument");
    int rc = check_adj_box(h->g);
    if (rc) return rc;
    const Geom& g = h->g;
    const int64_t n = g.n();
    if ((rc = ensure_adj(h, (size_t)n * (21 + 3 + 4)))) return rc;
    double* J = h->adj;
    double* JU = J + 21 * n;
    double* t1 = JU + 3 * n;
    double* t2 = t1 + n;
    double* la = t2 + n;
    double* lb = la + n;
    const hipStream_t s = S(stream);
    VLS_HIP(launch_advect_stage<double>(t1, F, nullptr, vx, vy, vz, dt, g, 0, s));
    VLS_HIP(launch_advect_stage<double>(t2, t1, F, vx, vy, vz, dt, g, 1, s));
    // stage 2: out = S2(t2; F)
    VLS_HIP(launch_adj_advect_jac(J, JU, t2, vx, vy, vz, dt, g, 2, s));
    VLS_HIP(launch_adj_transpose(la, J, lam_out, g, 0.0, s));
    VLS_HIP(launch_adj_velocity_grad(gx, gy, gz, JU, lam_out, n, s));
    VLS_HIP(launch_adj_axpby(lb, lam_out, stage_e(2), 0.0, n, s));
    // stage 1: t2 = S1(t1; F)
    VLS_HIP(launch_adj_advect_jac(J, JU, t1, vx, vy, vz, dt, g, 1, s));
    VLS_HIP(launch_adj_transpose(t2, J, la, g, 0.0, s));  // t2 <- adjoint of t1
    VLS_HIP(launch_adj_velocity_grad(gx, gy, gz, JU, la, n, s));
    VLS_HIP(launch_adj_axpby(lb, la, stage_e(1), 1.0, n, s));
    // stage 0: t1 = S0(F)
    VLS_HIP(launch_adj_advect_jac(J, JU, F, vx, vy, vz, dt, g, 0, s));
    VLS_HIP(launch_adj_transpose(lb, J, t2, g, 1.0, s));
    VLS_HIP(launch_adj_velocity_grad(gx, gy, gz, JU, t2, n, s));
    VLS_HIP(launch_copy<double>(lam_in, lb, n, s));
    return VLS_OK;
}

int vls_sdf3_tangent_advect_rk3(vls_sdf3_t h, const double* F, const double* Fdot, const double* vx, const double* vy,
                                const double* vz, const double* udx, const double* udy, const double* udz, double dt,
                                double* Fout, double* Fdot_out, vls_stream_t stream)
{
    VLS_REQUIRE(h && F && Fdot && vx && vy && vz && udx && udy && udz && Fout && Fdot_out, "null argument");
    VLS_REQUIRE(Fout != F && Fdot_out != Fdot, "outputs must not alias the inputs");
    int rc = check_adj_box(h->g);
    if (rc) return rc;
    const Geom& g = h->g;
    const int64_t n = g.n();
    if ((rc = ensure_adj(h, (size_t)n * (21 + 3 + 4)))) return rc;
    double* J = h->adj;
    double* JU = J + 21 * n;
    double* t1 = JU + 3 * n;
    double* t2 = t1 + n;
    double* d1 = t2 + n;
    double* d2 = d1 + n;
    const hipStream_t s = S(stream);
    // stage 0
    VLS_HIP(launch_adj_advect_jac(J, JU, F, vx, vy, vz, dt, g, 0, s));
    VLS_HIP(launch_adj_apply(d1, J, Fdot, g, s));
    VLS_HIP(launch_adj_velocity_tangent(d1, JU, udx, udy, udz, n, s));
    VLS_HIP(launch_advect_stage<double>(t1, F, nullptr, vx, vy, vz, dt, g, 0, s));
    // stage 1
    VLS_HIP(launch_adj_advect_jac(J, JU, t1, vx, vy, vz, dt, g, 1, s));
    VLS_HIP(launch_adj_apply(d2, J, d1, g, s));
    VLS_HIP(launch_adj_axpby(d2, Fdot, stage_e(1), 1.0, n, s));
    VLS_HIP(launch_adj_velocity_tangent(d2, JU, udx, udy, udz, n, s));
    VLS_HIP(launch_advect_stage<double>(t2, t1, F, vx, vy, vz, dt, g, 1, s));
    // stage 2
    VLS_HIP(launch_adj_advect_jac(J, JU, t2, vx, vy, vz, dt, g, 2, s));
    VLS_HIP(launch_adj_apply(Fdot_out, J, d2, g, s));
    VLS_HIP(launch_adj_axpby(Fdot_out, Fdot, stage_e(2), 1.0, n, s));
    VLS_HIP(launch_adj_velocity_tangent(Fdot_out, JU, udx, udy, udz, n, s));
    VLS_HIP(launch_advect_stage<double>(Fout, t2, F, vx, vy, vz, dt, g, 2, s));
    return VLS_OK;
}

// the reinitialization map R(F0): set_distance(F0), mask F0 < 0, `iterations`
// WENO5-RK3 iterations.  Recomputes the forward iterations, keeping every
// stage field, then sweeps back.  Leaves the handle with F0's distances.
int vls_sdf3_adjoint_reinit(vls_sdf3_t h, const double* F0, int iterations, const double* lam_out, double* lam_in,
                            vls_stream_t stream)
{
    VLS_REQUIRE(h && F0 && lam_out && lam_in, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    VLS_REQUIRE((const void*)F0 != (const void*)lam_in, "lam_in must not alias F0");
    int rc = check_adj_box(h->g);
    if (rc) return rc;
    const Geom& g = h->g;
    const int64_t n = g.n();
    const int K = iterations;
    if ((rc = ensure_adj(h, (size_t)n * (21 + 7 + 7 + 4 + 3 * (size_t)K + 1)))) return rc;
    double* J = h->adj;
    double* JP = J + 21 * n;
    double* lamP = JP + 7 * n;
    double* lam = lamP + 7 * n;
    double* lt0 = lam + n;
    double* lt1 = lt0 + n;
    double* lx = lt1 + n;
    double* X = lx + n;  // X[k] = X + k*n, k = 0..K; then T0[k], T1[k]
    double* T0 = X + (size_t)(K + 1) * n;
    double* T1 = T0 + (size_t)K * n;
    const hipStream_t s = S(stream);
    const DSet ms = main_set(h);
    if ((rc = set_distance_into<double>(h, 6, ms, h->min_dist, F0, (const double*)nullptr, false, s))) return rc;
    h->dense_ready = false;
    VLS_HIP(launch_set_mask_bit<double>(h->flags, F0, g, s));
    VLS_HIP(launch_copy<double>(X, F0, n, s));
    ReinitSeq rs;
    if ((rc = reinit_begin(h, rs, s))) return rc;
    for (int k = 0; k < K; ++k) {
        double* xk = X + (size_t)k * n;
        if ((rc = reinit_stage<double>(h, rs, T0 + (size_t)k * n, xk, nullptr, ms, 0, s))) return rc;
        if ((rc = reinit_stage<double>(h, rs, T1 + (size_t)k * n, T0 + (size_t)k * n, xk, ms, 1, s))) return rc;
        if ((rc = reinit_stage<double>(h, rs, xk + n, T1 + (size_t)k * n, xk, ms, 2, s))) return rc;
    }
    VLS_HIP(launch_copy<double>(lam, lam_out, n, s));
    VLS_HIP(launch_adj_fill(lamP, 0.0, 7 * n, s));
    for (int k = K - 1; k >= 0; --k) {
        const double* xk = X + (size_t)k * n;
        // stage 2: X[k+1] = S2(T1[k]; X[k])
        VLS_HIP(launch_adj_reinit_jac(J, JP, T1 + (size_t)k * n, h->flags, ms.lv, g, 2, s));
        VLS_HIP(launch_adj_transpose(lt1, J, lam, g, 0.0, s));
        VLS_HIP(launch_adj_param_grad(lamP, JP, lam, ms.lv, s));
        VLS_HIP(launch_adj_axpby(lx, lam, stage_e(2), 0.0, n, s));
        // stage 1: T1[k] = S1(T0[k]; X[k])
        VLS_HIP(launch_adj_reinit_jac(J, JP, T0 + (size_t)k * n, h->flags, ms.lv, g, 1, s));
        VLS_HIP(launch_adj_transpose(lt0, J, lt1, g, 0.0, s));
        VLS_HIP(launch_adj_param_grad(lamP, JP, lt1, ms.lv, s));
        VLS_HIP(launch_adj_axpby(lx, lt1, stage_e(1), 1.0, n, s));
        // stage 0: T0[k] = S0(X[k])
        VLS_HIP(launch_adj_reinit_jac(J, JP, xk, h->flags, ms.lv, g, 0, s));
        VLS_HIP(launch_adj_transpose(lx, J, lt0, g, 1.0, s));
        VLS_HIP(launch_adj_param_grad(lamP, JP, lt0, ms.lv, s));
        VLS_HIP(launch_copy<double>(lam, lx, n, s));
    }
    VLS_HIP(launch_copy<double>(lam_in, lam, n, s));
    VLS_HIP(launch_adj_dist(lam_in, F0, lamP, ms.lv, g, s));
    return VLS_OK;
}

int vls_sdf3_tangent_reinit(vls_sdf3_t h, const double* F0, const double* F0dot, int iterations, double* Fout,
                            double* Fdot_out, vls_stream_t stream)
{
    VLS_REQUIRE(h && F0 && F0dot && Fout && Fdot_out, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    VLS_REQUIRE(Fout != F0 && Fdot_out != F0dot && (const void*)Fout != (const void*)Fdot_out,
                "outputs must not alias the inputs");
    int rc = check_adj_box(h->g);
    if (rc) return rc;
    const Geom& g = h->g;
    const int64_t n = g.n();
    if ((rc = ensure_adj(h, (size_t)n * (21 + 7 + 7 + 6)))) return rc;
    double* J = h->adj;
    double* JP = J + 21 * n;
    double* Pd = JP + 7 * n;
    double* t0 = Pd + 7 * n;
    double* t1 = t0 + n;
    double* d0 = t1 + n;
    double* d1 = d0 + n;
    double* X = d1 + n;
    double* Xd = X + n;
    const hipStream_t s = S(stream);
    const DSet ms = main_set(h);
    if ((rc = set_distance_into<double>(h, 6, ms, h->min_dist, F0, (const double*)nullptr, false, s))) return rc;
    h->dense_ready = false;
    VLS_HIP(launch_set_mask_bit<double>(h->flags, F0, g, s));
    VLS_HIP(launch_adj_fill(Pd, 0.0, 7 * n, s));
    VLS_HIP(launch_tan_dist(Pd, F0, F0dot, ms.lv, g, s));
    VLS_HIP(launch_copy<double>(X, F0, n, s));
    VLS_HIP(launch_copy<double>(Xd, F0dot, n, s));
    ReinitSeq rs;
    if ((rc = reinit_begin(h, rs, s))) return rc;
    for (int k = 0; k < iterations; ++k) {
        VLS_HIP(launch_adj_reinit_jac(J, JP, X, h->flags, ms.lv, g, 0, s));
        VLS_HIP(launch_adj_apply(d0, J, Xd, g, s));
        VLS_HIP(launch_adj_param_tangent(d0, JP, Pd, ms.lv, s));
        if ((rc = reinit_stage<double>(h, rs, t0, X, nullptr, ms, 0, s))) return rc;
        VLS_HIP(launch_adj_reinit_jac(J, JP, t0, h->flags, ms.lv, g, 1, s));
        VLS_HIP(launch_adj_apply(d1, J, d0, g, s));
        VLS_HIP(launch_adj_axpby(d1, Xd, stage_e(1), 1.0, n, s));
        VLS_HIP(launch_adj_param_tangent(d1, JP, Pd, ms.lv, s));
        if ((rc = reinit_stage<double>(h, rs, t1, t0, X, ms, 1, s))) return rc;
        VLS_HIP(launch_adj_reinit_jac(J, JP, t1, h->flags, ms.lv, g, 2, s));
        VLS_HIP(launch_adj_apply(d0, J, d1, g, s));
        VLS_HIP(launch_adj_axpby(d0, Xd, stage_e(2), 1.0, n, s));
        VLS_HIP(launch_adj_param_tangent(d0, JP, Pd, ms.lv, s));
        if ((rc = reinit_stage<double>(h, rs, X, t1, X, ms, 2, s))) return rc;
        VLS_HIP(launch_copy<double>(Xd, d0, n, s));
    }
    VLS_HIP(launch_copy<double>(Fout, X, n, s));
    VLS_HIP(launch_copy<double>(Fdot_out, Xd, n, s));
    return VLS_OK;
}

}  // extern "C"
