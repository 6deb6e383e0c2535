// C-ABI layer: argument checks, the @SDF3 handle and the wrapper-level
// iteration loops.  Every loop issues its stage kernels back to back on the
// caller's stream; nothing here synchronizes except the calls that return a
// host scalar.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

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

int check_geom(int rows, int cols, int pges, double dx, double dy, double dz)
{
    if (rows <= 0 || cols <= 0 || pges <= 0) return fail(VLS_ERR_INVALID, "box %dx%dx%d is empty", rows, cols, pges);
    if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return fail(VLS_ERR_INVALID, "spacings must be positive");
    if ((int64_t)rows * cols * pges > ((int64_t)1 << 40)) return fail(VLS_ERR_INVALID, "box too large");
    if (pges > 65535 || (cols + 3) / 4 > 65535) return fail(VLS_ERR_INVALID, "box exceeds the launch grid");
    return VLS_OK;
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
        VLS_HIP(launch_extend_stage<T>(t0, C, nullptr, vx, vy, vz, ds.flags, ds.lv, h->lcp, h->lvel, g, 0, s));
        VLS_HIP(launch_extend_stage<T>(t1, t0, C, vx, vy, vz, ds.flags, ds.lv, h->lcp, h->lvel, g, 1, s));
        VLS_HIP(launch_extend_stage<T>(C, t1, C, vx, vy, vz, ds.flags, ds.lv, h->lcp, h->lvel, g, 2, s));
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

template <typename T>
int set_distance_into(vls_sdf3* h, int order, const DSet& ds, double* min_dist, const T* F, const T* mask_src,
                      hipStream_t s)
{
    const Geom& g = h->g;
    VLS_HIP(launch_fill_spacing(ds.d, g, ds.lv.count, s));
    VLS_HIP(launch_boundary_location<T>(order, ds.d, F, g, s));
    VLS_HIP(launch_min_dist_flags<T>(min_dist, ds.flags, constify(ds.d), mask_src, g, ds.lv.idx, ds.lv.count, s));
    VLS_HIP(launch_list_gather(ds.lv, constify(ds.d), g, s));
    return VLS_OK;
}

// ---- wrapper-level operations, for fields of T ------------------------------
template <typename T>
int sdf3_weno5_rk3_reinit(vls_sdf3* h, T* F, int iterations, hipStream_t s)
{
    VLS_REQUIRE(h && F, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    VLS_HIP(launch_set_mask_bit<T>(h->flags, F, h->g, s));
    return reinit_loop<T>(h, F, iterations, main_set(h), s);
}

template <typename T>
int sdf3_weno5_rk3_extend(vls_sdf3* h, const T* F, T* C, int iterations, hipStream_t s)
{
    VLS_REQUIRE(h && F && C, "null argument");
    VLS_REQUIRE(iterations >= 0, "negative iteration count");
    VLS_REQUIRE((const void*)F != (const void*)C, "C must not alias F");
    const Geom& g = h->g;
    const DSet ms = main_set(h);
    T *vx = vel<T>(h, 0), *vy = vel<T>(h, 1), *vz = vel<T>(h, 2);
    VLS_HIP(launch_upwind_normal<T>(vx, vy, vz, F, constify(h->d), g, s));
    VLS_HIP(launch_extension_velocity<T>(vx, vy, vz, F, 1e-6, g, s));
    VLS_HIP(launch_list_gather_vel<T>(h->lvel, ms.lv, vx, vy, vz, g, s));
    return extend_loop<T>(h, C, iterations, ms, s);
}

template <typename T>
int sdf3_surface_redistance(vls_sdf3* h, const T* F, T* A, int iter1, int iter2, hipStream_t s)
{
    VLS_REQUIRE(h && F && A, "null argument");
    VLS_REQUIRE(iter1 >= 0 && iter2 >= 0, "negative iteration count");
    VLS_REQUIRE((const void*)F != (const void*)A, "A must not alias F");
    const Geom& g = h->g;
    int rc;
    // A's distances, deltat and mask (WENO5RK3ClosetPointSurfaceRedistance.m:8-25)
    const DSet as = aux_set(h), fs = f_set(h);
    if ((rc = set_distance_into<T>(h, 6, as, nullptr, A, A, s))) return rc;
    // F's distances and the extension velocity with floor 1e-14 (:30-74)
    if ((rc = set_distance_into<T>(h, 6, fs, nullptr, F, nullptr, s))) return rc;
    T *vx = vel<T>(h, 0), *vy = vel<T>(h, 1), *vz = vel<T>(h, 2);
    VLS_HIP(launch_upwind_normal<T>(vx, vy, vz, F, constify(h->fd), g, s));
    VLS_HIP(launch_extension_velocity<T>(vx, vy, vz, F, 1e-14, g, s));
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
    for (int it = 0; it < steps; ++it) {
        VLS_HIP(launch_advect_stage<T>(t0, F, nullptr, vx, vy, vz, dt, g, 0, s));
        VLS_HIP(launch_advect_stage<T>(t1, t0, F, vx, vy, vz, dt, g, 1, s));
        VLS_HIP(launch_advect_stage<T>(F, t1, F, vx, vy, vz, dt, g, 2, s));
    }
    return VLS_OK;
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

}  // namespace

extern "C" {

const char* vls_last_error(void) { return g_err.c_str(); }
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
    const size_t bytes = sizeof(double) * ((size_t)n * kFields + (size_t)reduce_blocks() * 4 + 8);
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
    if (hipMemset(h->int_pool, 0, sizeof(int) * ((size_t)n * 4 + 16)) != hipSuccess) {
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
    h->partial = p; p += (size_t)reduce_blocks() * 4;
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

int vls_sdf3_destroy(vls_sdf3_t h)
{
    if (!h) return VLS_OK;
    hipError_t e1 = hipFree(h->pool), e2 = hipFree(h->flag_pool), e3 = hipFree(h->int_pool);
    delete h;
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(VLS_ERR_HIP, "hipFree failed");
    return VLS_OK;
}

int vls_sdf3_set_distance(vls_sdf3_t h, const double* F, vls_stream_t stream)
{
    VLS_REQUIRE(h && F, "null argument");
    return set_distance_into<double>(h, 6, main_set(h), h->min_dist, F, nullptr, S(stream));
}

int vls_sdf3_set_distance_f32(vls_sdf3_t h, const float* F, vls_stream_t stream)
{
    VLS_REQUIRE(h && F, "null argument");
    return set_distance_into<float>(h, 6, main_set(h), h->min_dist, F, nullptr, S(stream));
}

int vls_sdf3_distances(vls_sdf3_t h, const double** out7)
{
    VLS_REQUIRE(h && out7, "null argument");
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

int vls_sdf3_eno2_reinit(vls_sdf3_t h, double* F, int iterations, int rk2, double cfl, vls_stream_t stream)
{
    VLS_REQUIRE(h && F, "null argument");
    VLS_REQUIRE(iterations >= 0 && cfl > 0, "bad iterations/cfl");
    const hipStream_t s = S(stream);
    const Geom& g = h->g;
    // ENO2 distances of the argument and mask = F<0 (ENORK2Reinitialization.m:7-25)
    int rc = set_distance_into<double>(h, 2, aux_set(h), nullptr, F, F, s);
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
