// Internal interface between the kernels (vls_kernels.hip) and the C-ABI
// layer (vls_api.hip).  Not installed; the public interface is
// include/vesicle_ls.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vls {

struct Box;

// six per-axis arrays: r[a] = right/front/up distance (xpr, ypf, zpu),
// l[a] = left/back/down distance (xpl, ypb, zpd); a = 0 x(cols), 1 y(rows), 2 z(pges)
struct Dist6 {
    double* r[3];
    double* l[3];
};
struct Dist6c {
    const double* r[3];
    const double* l[3];
};
struct Fields18 {
    double* f[18];
};
// the auxiliary-field toolbox: 27 outputs and the 13 toolbox fields of F it reads
struct Fields27 {
    double* f[27];
};
struct Fields13c {
    const double* f[13];
};

// per-point flag bits of the fused stage kernels
enum : uint8_t {
    FLAG_MASK = 1,     // phi0 < 0 at entry of the reinitialization
    FLAG_SUBCELL = 2,  // some subcell distance differs from the grid spacing
};

inline Dist6c constify(const Dist6& d)
{
    return Dist6c{{d.r[0], d.r[1], d.r[2]}, {d.l[0], d.l[1], d.l[2]}};
}

// ---- host launchers (all asynchronous on `s`) -------------------------------
// an interface list: point indices (count on the device) and, in list order,
// the six distances (rows of `cap` doubles: xpr,xpl,ypf,ypb,zpu,zpd) and the
// pseudo-time step (0.3*min_dist, -1 when frozen)
struct ListView {
    int* idx;
    int* count;
    double* dist;
    double* dt;
    int cap;
};


// interface schemes of the fused stage kernels and the list passes: the
// nonuniform WENO of the WENO5 path, the cubic ENO of the WENO-RK3 variants,
// the quadratic ENO of the ENO-RK2 extension (list passes of the batch only)
enum { IFACE_WENO = 0, IFACE_CUBIC = 1, IFACE_ENO2 = 2 };

struct Geom {
    int rows, cols, pges;
    double dx, dy, dz;
    int tr = 64;      // rows per tile of the bulk stage kernels: 64 or 32
    int zd = 0;       // pages per block of the z-marching bulk stage kernels (0: by the box size)
    int64_t n() const { return (int64_t)rows * cols * pges; }
};

// Launchers templated on T take fields of T = double or float (FP32 bulk: the
// stage kernels' bulk arithmetic is FP32; subcell distances, interface-list
// arithmetic and reductions stay FP64).
template <typename T>
hipError_t launch_boundary_location(int order, const Dist6& d, const T* lsf, const Geom& g, hipStream_t s);
hipError_t launch_zero_ints(int* a, int n, hipStream_t s);
// setDistance: the flags (FLAG_SUBCELL where a distance differs from the
// spacing, FLAG_MASK where mask_src < 0; mask_src may be null), the
// interface list with its list-order distances and time steps (lv.idx null:
// no list), and, if d.r[0] is set, the dense distance arrays and min_dist
// (may be null).  cand (n ints) and *cand_count are work memory.  Four
// launches: zero the counts, a pass over every point, a pass over the points
// with a crossing.
template <typename T>
hipError_t launch_set_distance(int order, const Dist6& d, double* min_dist, uint8_t* flags, const T* mask_src,
                               const T* lsf, const ListView& lv, int* cand, int* cand_count, const Geom& g,
                               hipStream_t s);
// the dense arrays (six distances, min_dist) of a distance set located
// without them, from its interface list
hipError_t launch_fill_dense(const Dist6& d, double* min_dist, const ListView& lv, const Geom& g, hipStream_t s);
int iface_blocks(const Geom& g);
// pages per block of the z-marching bulk stage kernels (g.zd, or chosen by the box size)
int zmarch_depth(const Geom& g);
template <typename T>
hipError_t launch_set_mask_bit(uint8_t* flags, const T* F, const Geom& g, hipStream_t s);
hipError_t launch_re_step(double* step, const double* lsf, const uint8_t* mask, const double* deltat,
                          const Dist6c& d, const Geom& g, hipStream_t s);
hipError_t launch_eno2_re_step(double* step, const double* lsf, const uint8_t* mask, const double* deltat,
                               const Dist6c& d, const Geom& g, hipStream_t s);
// one WENO5-RK3 reinitialization stage (two launches); `seq` numbers the
// stages of one reinitialization call from 0, over_count holds 2 ints that
// are zero before seq 0
template <typename T>
hipError_t launch_reinit_stage(T* out, const T* fs, const T* fn, const uint8_t* flags, const ListView& lv,
                               const Geom& g, int stage, int* over, int* over_count, int seq, hipStream_t s);
// interface values of C at the listed points, list order (6 rows of cap)
template <typename T>
hipError_t launch_list_interp(double* lcp, const ListView& lv, const T* C, const Geom& g, hipStream_t s);
// velocity at the listed points, list order (3 rows of cap)
template <typename T>
hipError_t launch_list_gather_vel(double* lvel, const ListView& lv, T* vx, const T* vy, const T* vz,
                                  const Geom& g, hipStream_t s);
hipError_t launch_eno2_stage(double* out, const double* fs, const double* fn, const uint8_t* flags,
                             const Dist6c& d, const Geom& g, double cfl, int heun2, hipStream_t s);
template <typename T>
hipError_t launch_upwind_normal(T* nx, T* ny, T* nz, const T* lsf, const Dist6c& d, const Geom& g, hipStream_t s);
// upwind normal and extension velocity of F (the two below, fused): the
// points without FLAG_SUBCELL, then the listed ones (distances from the list)
template <typename T>
hipError_t launch_extension_normal(T* vx, T* vy, T* vz, const T* F, const uint8_t* flags, const ListView& lv,
                                   double floor_, const Geom& g, hipStream_t s);
template <typename T>
hipError_t launch_extension_velocity(T* vx, T* vy, T* vz, const T* F, double floor_, const Geom& g, hipStream_t s);

// WENO-RK3 / ENO-RK2 variants (vls_variants.hip), FP64.  Drop-in kernels:
hipError_t launch_wenork3_re_step(double* step, const double* lsf, const uint8_t* mask, const double* deltat,
                                  const Dist6c& d, const Geom& g, hipStream_t s);
hipError_t launch_enork2_upwind_normal(double* nx, double* ny, double* nz, const double* lsf, const Dist6c& d,
                                       const Geom& g, hipStream_t s);
hipError_t launch_enork2_boundary_interpolate(const Dist6& cp, const Dist6c& d, const double* lsf, const Geom& g,
                                              hipStream_t s);
hipError_t launch_enork2_extend_step(double* step, const double* deltat, const double* lsf, const double* vx,
                                     const double* vy, const double* vz, const Dist6c& d, const Dist6c& cp,
                                     const Geom& g, hipStream_t s);
hipError_t launch_wenork3_upwind_normal(double* nx, double* ny, double* nz, const double* lsf, const Dist6c& d,
                                        const Geom& g, hipStream_t s);
hipError_t launch_wenork3_boundary_interpolate(const Dist6& cp, const Dist6c& d, const double* lsf, const Geom& g,
                                               hipStream_t s);
hipError_t launch_wenork3_extend_step(double* step, const double* deltat, const double* lsf, const double* vx,
                                      const double* vy, const double* vz, const Dist6c& d, const Dist6c& cp,
                                      const Geom& g, hipStream_t s);
// the WENO-RK3 variants on the fused stage kernels of the WENO5 path (same
// z-marching bulk; cubic ENO instead of the nonuniform WENO at the listed and
// deferred points), FP64; lcp from launch_list_interp_cubic
hipError_t launch_cubic_reinit_stage(double* out, const double* fs, const double* fn, const uint8_t* flags,
                                     const ListView& lv, const Geom& g, int stage, int* over, int* over_count,
                                     int seq, hipStream_t s);
hipError_t launch_cubic_extend_stage(double* out, const double* cs, const double* cn, const double* vx,
                                     const double* vy, const double* vz, const ListView& lv, const double* lcp,
                                     const double* lvel, const Geom& g, int stage, hipStream_t s);
hipError_t launch_list_interp_cubic(double* lcp, const ListView& lv, const double* C, const Geom& g, hipStream_t s);
// fused stage of the ENO-RK2 extension loop (operator + Heun combination;
// dt = 0.3*min of the six distances d; cp the interface values; d and cp are
// read only where flags has FLAG_SUBCELL)
hipError_t launch_enork2_extend_stage(double* out, const double* cs, const double* cn, const double* vx,
                                      const double* vy, const double* vz, const uint8_t* flags, const Dist6c& d,
                                      const Dist6c& cp, const Geom& g, int heun2, hipStream_t s);
hipError_t launch_boundary_interpolant(const Dist6& cp, const Dist6c& d, const double* C, const Geom& g,
                                       hipStream_t s);
hipError_t launch_extend_step(double* step, const double* deltat, const double* C, const double* vx,
                              const double* vy, const double* vz, const Dist6c& d, const Dist6c& cp,
                              const Geom& g, hipStream_t s);
template <typename T>
hipError_t launch_extend_stage(T* out, const T* cs, const T* cn, const T* vx, const T* vy, const T* vz,
                               const ListView& lv, const double* lcp, const double* lvel, const Geom& g, int stage,
                               hipStream_t s);
hipError_t launch_advection_step(double* astep, const double* vx, const double* vy, const double* vz,
                                 const double* lsf, const Geom& g, hipStream_t s);
template <typename T>
hipError_t launch_advect_stage(T* out, const T* fs, const T* fn, const T* vx, const T* vy, const T* vz, double dt,
                               const Geom& g, int stage, hipStream_t s);
hipError_t launch_weno_derivative(const Dist6& out, const double* lsf, const Geom& g, hipStream_t s);
// auxi_set_calculus_toolbox (4_0_Calculus_ToolBox/auxi_set_calculus_toolbox.cu:9-199): out
// = Ax..ADiracDelta of the auxiliary field A (lsf) on the surface of F, whose
// toolbox fields Fx, Fy, Fz, FGradMag, Nx, Ny, Nz, Fxx, Fyy, Fzz, Fxy, Fyz, Fzx
// are fin; clamped neighbours
hipError_t launch_auxi_set_calculus_toolbox(const Fields27& o, const double* lsf, const double* ahp,
                                            const Fields13c& fin, const Geom& g, double ds, hipStream_t s);
// the MATLAB tail of GPUAsetCalculusToolBox.m:55-62 with the periodic central
// differences of @GD3 (GD3.m:112-132 index maps): AFDiracDelta and ADiracDeltaDn
hipError_t launch_aux_toolbox_tail(double* afdd, double* addn, const double* ah, const double* h, const double* a,
                                   const double* cx, const double* cy, const double* cz, const double* ncross,
                                   const double* add, const double* agrad, const Geom& g, hipStream_t s);
hipError_t launch_positive_part(double* out, const double* a, int64_t n, hipStream_t s);

// vesicle relaxation step (vls_relax.hip): the @GD3 smoothers through rocFFT
// (exact = 0: spectral symbol of smoothFFT; 1: the finite-difference system of
// smoothGMRES solved exactly), c = alpha*dt; the operator itself; the bending
// speed; surface integrals, Dt and the tension/pressure solve with the normal
// speed (partial: reduce_blocks()*13 doubles; out10 on the device)
hipError_t launch_smooth(double* out, const double* rhs, const Geom& g, double c, int diffusion, int exact,
                         hipStream_t s);
hipError_t launch_smooth_operator(double* y, const double* x, const Geom& g, double c, int diffusion, hipStream_t s);
hipError_t launch_bending_speed(double* out, const double* H, const double* G, const Geom& g, double kappa,
                                double sponc, hipStream_t s);
hipError_t launch_tension_pressure(double* out10, double* partial, double* speed, const double* F, const double* H,
                                   const double* Vb, const double* dirac, const double* gmag, const double* heav,
                                   const Geom& g, double ds, double a0, double v0, double dt, double cfl,
                                   hipStream_t s);
hipError_t launch_set_calculus_toolbox(const Fields18& o, const double* lsf, const double* hp, const Geom& g,
                                       double ds, hipStream_t s);
// reductions: `partial` holds at least reduce_blocks()*4 doubles, `result` 2
int reduce_blocks();
// result[0..1] = (area, volume) sums times `scale`
template <typename T>
hipError_t launch_area_volume_cosine(double* result, double* partial, const T* F, const Geom& g, double eps,
                                     double scale, hipStream_t s);
template <typename T>
hipError_t launch_area_volume_towers(double* result, double* partial, const T* F, const Geom& g, double ds,
                                     double scale, hipStream_t s);
template <typename T>
hipError_t launch_normal_velocity(T* out, const T* F, const T* vx, const T* vy, const T* vz, const Geom& g,
                                  hipStream_t s);
template <typename T>
hipError_t launch_max_abs3(double* result, double* partial, const T* a, const T* b, const T* c, const Geom& g,
                           hipStream_t s);
template <typename T>
hipError_t launch_copy(T* dst, const T* src, int64_t n, hipStream_t s);
hipError_t launch_widen(double* dst, const float* src, int64_t n, hipStream_t s);
hipError_t launch_narrow(float* dst, const double* src, int64_t n, hipStream_t s);

// ---- batches of grids (vls_batch.hip) -----------------------------------------
// One grid of a batch: its box, its first element in the packed fields (every
// field of a batch is the grids' column-major arrays back to back) and its
// first tile in the batch's tile list (tiles of tr rows x 256/tr cols x 1
// page, page-major).
// per-axis constants of the Jiang-Peng form of the uniform WENO5 derivative
// (vls_device.hpp): k = 1/(12 ds), e = 12/13 * 1e-6 * ds^2, in FP64 (the
// expressions the CPU oracle evaluates); the FP32 form takes its indicators
// divided by 16/13, so its floor is e * 13/16, rounded to FP32 with k
struct JPd {
    double k, e;
};
struct JPf {
    float k, e;
};

__host__ __device__ inline JPd jpd(double ds) { return JPd{1.0 / (12.0 * ds), 12.0 / 13.0 * 1e-6 * ds * ds}; }

__host__ __device__ inline void set_jp(JPd jd[3], JPf jf[3], double dx, double dy, double dz)
{
    const double sp[3] = {dx, dy, dz};
    for (int a = 0; a < 3; ++a) {
        jd[a] = jpd(sp[a]);
        jf[a].k = (float)jd[a].k;
        jf[a].e = (float)(jd[a].e * (13.0 / 16.0));
    }
}

struct BGrid {
    int rows, cols, pges;
    int ntiles;       // tiles of this grid
    int tile0;        // first tile of this grid
    int64_t off;      // first element of this grid
    double dx, dy, dz;
    double inv[3];    // 1/dx, 1/dy, 1/dz
    double dt_bulk;   // 0.3*min(dx, dy, dz)
    JPd jd[3];        // the WENO5 form's per-axis constants (set_jp)
    JPf jf[3];
};

// device view of a batch
// one block's share of a z-marching batch launch: tile column (r0, c0) of
// grid g over pages [p0, p1)
struct ZItem {
    int g, r0, c0, p0, p1;
};

struct BatchView {
    const BGrid* grids;    // ngrids descriptors
    const int* tile_grid;  // grid of each tile
    const int64_t* offs;   // ngrids + 1 grid offsets (offs[ngrids] = total points)
    int ngrids, ntiles;
    int blocks;            // persistent blocks of a tile-loop launch (multiple of 8)
    int tr;                // rows per tile: 64 or 32
    const ZItem* items;    // z-march work list of the bulk stage kernels
    int nitems;
    int* list_grid;        // grid of each interface-list entry (written with the list)
};

// tiles of one grid with `tr` rows per tile (tr x 256/tr columns)
__host__ __device__ inline int grid_tiles(int rows, int cols, int pges, int tr)
{
    const int tc = 256 / tr;
    return ((rows + tr - 1) / tr) * ((cols + tc - 1) / tc) * pges;
}

// list building: subcell distances (the locator of `order`: 6, 2 for the
// quadratic ENO, 3 for the cubic), flags (mask bit from mask_src < 0 when
// mask_src is not null) and the interface list of every grid, in two passes;
// cand (total ints) and *cand_count are work memory; *lv.count and
// *cand_count must be zero on entry
template <typename T>
hipError_t launch_batch_build_list(const BatchView& B, const T* F, const T* mask_src, uint8_t* flags,
                                   const ListView& lv, int order, int* cand, int* cand_count, hipStream_t s);
template <typename T>
hipError_t launch_batch_set_mask_bit(uint8_t* flags, const T* F, int64_t n, hipStream_t s);
// `scheme` of the listed and deferred points: 0 the nonuniform WENO of the
// WENO5 path, 1 the cubic ENO of the WENO-RK3 variants (FP64 only)
template <typename T>
hipError_t launch_batch_reinit_stage(const BatchView& B, T* out, const T* fs, const T* fn, const uint8_t* flags,
                                     const ListView& lv, int stage, int* over, int* over_count, int seq,
                                     hipStream_t s, int scheme = 0);
template <typename T>
hipError_t launch_batch_extend_stage(const BatchView& B, T* out, const T* cs, const T* cn, const T* vx,
                                     const T* vy, const T* vz, const ListView& lv, const double* lcp,
                                     const double* lvel, int stage, hipStream_t s,
                                     int scheme = 0);
// one ENO-RK2 extension stage of every grid (heun2: 0 first, 1 second Heun
// stage); lcp: the ENO2 interface values of the listed points
hipError_t launch_batch_eno2_extend_stage(const BatchView& B, double* out, const double* cs, const double* cn,
                                          const double* vx, const double* vy, const double* vz, const uint8_t* flags,
                                          const ListView& lv, const double* lcp, int heun2, hipStream_t s);
// dt: one step per grid (device)
template <typename T>
hipError_t launch_batch_advect_stage(const BatchView& B, T* out, const T* fs, const T* fn, const T* vx,
                                     const T* vy, const T* vz, const double* dt, int stage, hipStream_t s);
// upwind normal of F (subcell-aware) and v = sign(F) n/max(|n|, floor);
// scheme 0: WENO5 (upwind_normal.cu), 1: the cubic ENO of WENORK3Extend,
// 2: the ENO2 of ENORK2Extend (1 and 2 FP64 only)
template <typename T>
hipError_t launch_batch_extension_velocity(const BatchView& B, T* vx, T* vy, T* vz, const T* F,
                                           const uint8_t* flags, const ListView& lv, double floor_, hipStream_t s,
                                           int scheme = 0);
// interface values of C at the listed points (scheme as above: the quintic,
// the cubic or the quadratic ENO interpolant)
template <typename T>
hipError_t launch_batch_list_interp(const BatchView& B, double* lcp, const ListView& lv, const T* C, hipStream_t s,
                                    int scheme = 0);
template <typename T>
hipError_t launch_batch_list_gather_vel(double* lvel, const ListView& lv, T* vx, const T* vy, const T* vz,
                                        hipStream_t s);
template <typename T>
hipError_t launch_batch_normal_velocity(const BatchView& B, T* out, const T* F, const T* vx, const T* vy,
                                        const T* vz, hipStream_t s);
// per-grid cosine area and volume (eps = eps_factor*dx of each grid):
// result[2g], result[2g+1]; partial holds 4*ntiles doubles
template <typename T>
hipError_t launch_batch_area_volume_cosine(const BatchView& B, double* result, double* partial, const T* F,
                                           double eps_factor, hipStream_t s);
// per-grid step dt[g] = cfl*dx/max|u| (partial: ntiles doubles)
template <typename T>
hipError_t launch_batch_cfl_dt(const BatchView& B, double* dt, double* partial, const T* vx, const T* vy,
                               const T* vz, double cfl, hipStream_t s);
// synthetic inputs: u = curl of sum_m a_m cos(k_m.x + psi_m) per grid (modes:
// ngrids x nmodes x 7 doubles k(3), k x a (3), psi), and the ellipsoid
// sqrt(x^2/a^2+y^2/b^2+z^2/c^2)-1 (axes: 3 per grid); lo: 3 per grid
template <typename T>
hipError_t launch_batch_curl_velocity(const BatchView& B, T* vx, T* vy, T* vz, const double* modes, int nmodes,
                                      const double* lo, hipStream_t s);
template <typename T>
hipError_t launch_batch_ellipsoid(const BatchView& B, T* F, const double* axes, const double* lo, hipStream_t s);

// ---- discrete adjoint and tangent (vls_adjoint.hip), FP64, one grid ----------
// stage Jacobians: J = 21 rows of n (d out_i / d fs along the three lines
// through i, row a*7 + offset+3), JU = 3 rows of n (d out_i / d u_a,i),
// JP = 7 rows of lv.cap (d out / d (xpr..zpd, dt) of the listed points)
hipError_t launch_adj_advect_jac(double* J, double* JU, const double* fs, const double* vx, const double* vy,
                                 const double* vz, double dt, const Geom& g, int stage, hipStream_t s);
hipError_t launch_adj_reinit_jac(double* J, double* JP, const double* fs, const uint8_t* flags, const ListView& lv,
                                 const Geom& g, int stage, hipStream_t s);
// lam_in = beta*lam_in + J^T lam;  out = J x
hipError_t launch_adj_transpose(double* lam_in, const double* J, const double* lam, const Geom& g, double beta,
                                hipStream_t s);
hipError_t launch_adj_apply(double* out, const double* J, const double* x, const Geom& g, hipStream_t s);
hipError_t launch_adj_fill(double* y, double v, int64_t n, hipStream_t s);
hipError_t launch_adj_axpby(double* y, const double* x, double alpha, double beta, int64_t n, hipStream_t s);
hipError_t launch_adj_velocity_grad(double* gx, double* gy, double* gz, const double* JU, const double* lam,
                                    int64_t n, hipStream_t s);
hipError_t launch_adj_velocity_tangent(double* out, const double* JU, const double* ux, const double* uy,
                                       const double* uz, int64_t n, hipStream_t s);
hipError_t launch_adj_param_grad(double* lamP, const double* JP, const double* lam, const ListView& lv,
                                 hipStream_t s);
hipError_t launch_adj_param_tangent(double* out, const double* JP, const double* Pdot, const ListView& lv,
                                    hipStream_t s);
// distances of the listed points as functions of F0: adjoint (atomic adds
// into lamF0) and tangent (Pdot, 7 rows of lv.cap)
hipError_t launch_adj_dist(double* lamF0, const double* F0, const double* lamP, const ListView& lv, const Geom& g,
                           hipStream_t s);
hipError_t launch_tan_dist(double* Pdot, const double* F0, const double* F0dot, const ListView& lv, const Geom& g,
                           hipStream_t s);

}  // namespace vls
