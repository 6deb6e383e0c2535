// HIP kernels of the WENO-RK3 and ENO-RK2 variants of the reinitialization
// and the extension that the vesicle drivers also call
// (1_1_WENORK3_Reinitialization/weno_rk3_reinitialization.cu,
// 2_0_ENORK2_Extend/enork2_extend.cu, 2_1_WENORK3_Extend/wenork3_extend.cu;
// paths relative to src/7_0_paper/2_0_Vesicle/CUDA_Code/).
//
// Same layout and launch geometry as vls_kernels.hip: one thread per point,
// 64 rows x 4 cols of one page per block, clamped neighbours.  Two families:
//   * drop-in kernels with the argument meaning of the reference's entry
//     points (re_step, upwind_normal, boundary_interpolate, extend_step);
//   * the fused stage of the ENO-RK2 extension loop: the operator and its
//     Heun combination in one pass, the pseudo-time step 0.3*min of the six
//     distances derived in the kernel (ENORK2Extend.m:61-64).
// The WENO-RK3 wrapper loops run on the fused stage kernels of the WENO5 path
// (vls_kernels.hip) with the cubic-ENO interface scheme.
// The cubic crossing distances (cubic boundary_correction) are the ORDER 3
// case of k_boundary_location in vls_kernels.hip.
#include "vls_tile.hpp"

namespace vls {

namespace {

inline Box vbox(const Geom& g) { return Box{g.rows, g.cols, g.pges, g.dx, g.dy, g.dz}; }

__device__ __forceinline__ void point_dists(Dists& q, const Dist6c& d, idx_t ind)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) { q.r[a] = d.r[a][ind]; q.l[a] = d.l[a][ind]; }
}

// the fused stages read the six distances (and interface values) only at a
// point whose FLAG_SUBCELL says some distance differs from the spacing; every
// other point has all six equal to the spacings, which is what it gets here
__device__ __forceinline__ void point_dists(Dists& q, const Dist6c& d, idx_t ind, bool sub, const Box& b)
{
    if (sub) {
        point_dists(q, d, ind);
    } else {
        q.r[0] = q.l[0] = b.dx;
        q.r[1] = q.l[1] = b.dy;
        q.r[2] = q.l[2] = b.dz;
    }
}

// ---------------------------------------------------------------------------
// WENO-RK3 reinitialization step at a point (weno_rk3_reinitialization.cu:295-407):
// frozen next to the interface (any distance below 1e-6 dx), else the Godunov
// step of the cubic-ENO / WENO5 derivative pairs
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wenork3_reinit_residual(const double* __restrict__ f, int r, int c, int p,
                                                          const Box& b, const Dists& q, bool inside, double dt)
{
    if (frozen(q, b.dx)) return 0.0;
    const double sp[3] = {b.dx, b.dy, b.dz};
    double fo[3], bk[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7];
        line7(f, r, c, p, a, b, v);
        const bool cross = v[2] * v[3] < 0 || v[3] * v[4] < 0;
        wenork3_pair(fo[a], bk[a], v, q.r[a], q.l[a], sp[a], jpd(sp[a]), cross, 0.0, 0.0);
    }
    return godunov_step(inside, fo[0], bk[0], fo[1], bk[1], fo[2], bk[2], dt);
}

// WENO-RK3 extension step at a point (wenork3_extend.cu:372-495): frozen next
// to the interface, else cubic ENO with the interface values cp where an axis
// has a subcell distance, WENO5 elsewhere
__device__ __forceinline__ double wenork3_extend_residual(const double* __restrict__ f, int r, int c, int p,
                                                          const Box& b, const Dists& q, const Dist6c& cp, idx_t ind,
                                                          double vx, double vy, double vz, double dt)
{
    if (frozen(q, b.dx)) return 0.0;
    const double sp[3] = {b.dx, b.dy, b.dz};
    double fo[3], bk[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7];
        line7(f, r, c, p, a, b, v);
        const bool cross = q.r[a] < sp[a] || q.l[a] < sp[a];
        // the interface values are read where the axis crosses (they are 0
        // elsewhere, and unused)
        const double hf = cross ? cp.r[a][ind] : 0.0, hb = cross ? cp.l[a][ind] : 0.0;
        wenork3_pair(fo[a], bk[a], v, q.r[a], q.l[a], sp[a], jpd(sp[a]), cross, hf, hb);
    }
    return upwind_sum(fo, bk, vx, vy, vz, dt);
}

// ENO-RK2 extension step at a point (enork2_extend.cu:213-266; no frozen test)
// with the dense interface values cp
__device__ __forceinline__ double enork2_extend_residual(const double* __restrict__ f, int r, int c, int p,
                                                         const Box& b, const Dists& q, const Dist6c& cp, idx_t ind,
                                                         double vx, double vy, double vz, double dt)
{
    const double sp[3] = {b.dx, b.dy, b.dz};
    double hf[3], hb[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        hf[a] = (q.r[a] != sp[a]) ? cp.r[a][ind] : 0.0;
        hb[a] = (q.l[a] != sp[a]) ? cp.l[a][ind] : 0.0;
    }
    return enork2_residual(f, r, c, p, b, q, hf, hb, vx, vy, vz, dt);
}

}  // namespace

// ---------------------------------------------------------------------------
// drop-in kernels
// ---------------------------------------------------------------------------

// re_step: weno_rk3_reinitialization.cu:295-407
__global__ void __launch_bounds__(256) k_wenork3_re_step(double* __restrict__ step, const double* __restrict__ lsf,
                                                         const uint8_t* __restrict__ mask,
                                                         const double* __restrict__ deltat, Dist6c d, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    Dists q;
    point_dists(q, d, ind);
    step[ind] = wenork3_reinit_residual(lsf, r, c, p, b, q, mask[ind] != 0, deltat[ind]);
}

// upwind_normal: enork2_extend.cu:102-138 (upwind_normal_point :82-99: the ENO2
// derivative on the side whose node value, 0 at a crossing, is smaller)
__global__ void __launch_bounds__(256) k_enork2_upwind_normal(double* __restrict__ nx, double* __restrict__ ny,
                                                              double* __restrict__ nz, const double* __restrict__ lsf,
                                                              Dist6c d, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double sp[3] = {b.dx, b.dy, b.dz};
    double* out[3] = {nx, ny, nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7], sR, sL;
        line7(lsf, r, c, p, a, b, v);
        const double pr = d.r[a][ind], pl = d.l[a][ind];
        eno2_pair(sR, sL, v, pr, pl, sp[a]);
        const double vr = (pr == sp[a]) ? v[4] : 0.0, vl = (pl == sp[a]) ? v[2] : 0.0;
        out[a][ind] = (fabs(vr) < fabs(vl)) ? sR : sL;
    }
}

// boundary_interpolate: enork2_extend.cu:174-210
__global__ void __launch_bounds__(256) k_enork2_boundary_interpolate(Dist6 cp, Dist6c d,
                                                                     const double* __restrict__ lsf, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double sp[3] = {b.dx, b.dy, b.dz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7], cf, cb;
        line7(lsf, r, c, p, a, b, v);
        eno2_interface_values(cf, cb, d.r[a][ind], d.l[a][ind], sp[a], v);
        cp.r[a][ind] = cf;
        cp.l[a][ind] = cb;
    }
}

// extend_step: enork2_extend.cu:213-266
__global__ void __launch_bounds__(256) k_enork2_extend_step(double* __restrict__ step,
                                                            const double* __restrict__ deltat,
                                                            const double* __restrict__ lsf,
                                                            const double* __restrict__ vx,
                                                            const double* __restrict__ vy,
                                                            const double* __restrict__ vz, Dist6c d, Dist6c cp, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    Dists q;
    point_dists(q, d, ind);
    step[ind] = enork2_extend_residual(lsf, r, c, p, b, q, cp, ind, vx[ind], vy[ind], vz[ind], deltat[ind]);
}

// upwind_normal: wenork3_extend.cu:202-292 (upwind_normal_point :193-199:
// interface values 0; the derivative toward the smaller |p| neighbour)
__global__ void __launch_bounds__(256) k_wenork3_upwind_normal(double* __restrict__ nx, double* __restrict__ ny,
                                                               double* __restrict__ nz,
                                                               const double* __restrict__ lsf, Dist6c d, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double sp[3] = {b.dx, b.dy, b.dz};
    double* out[3] = {nx, ny, nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7], fo, bk;
        line7(lsf, r, c, p, a, b, v);
        const double r1 = d.r[a][ind], l1 = d.l[a][ind];
        wenork3_pair(fo, bk, v, r1, l1, sp[a], jpd(sp[a]), r1 < sp[a] || l1 < sp[a], 0.0, 0.0);
        out[a][ind] = (fabs(v[4]) < fabs(v[2])) ? fo : bk;
    }
}

// boundary_interpolate: wenork3_extend.cu:331-368
__global__ void __launch_bounds__(256) k_wenork3_boundary_interpolate(Dist6 cp, Dist6c d,
                                                                      const double* __restrict__ lsf, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double sp[3] = {b.dx, b.dy, b.dz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7], cf, cb;
        line7(lsf, r, c, p, a, b, v);
        cubic_interface_values(cf, cb, d.r[a][ind], d.l[a][ind], sp[a], v);
        cp.r[a][ind] = cf;
        cp.l[a][ind] = cb;
    }
}

// extend_step: wenork3_extend.cu:372-495
__global__ void __launch_bounds__(256) k_wenork3_extend_step(double* __restrict__ step,
                                                             const double* __restrict__ deltat,
                                                             const double* __restrict__ lsf,
                                                             const double* __restrict__ vx,
                                                             const double* __restrict__ vy,
                                                             const double* __restrict__ vz, Dist6c d, Dist6c cp,
                                                             Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    Dists q;
    point_dists(q, d, ind);
    step[ind] = wenork3_extend_residual(lsf, r, c, p, b, q, cp, ind, vx[ind], vy[ind], vz[ind], deltat[ind]);
}

// ---------------------------------------------------------------------------
// fused stage kernels of the wrapper loops
// ---------------------------------------------------------------------------

// ENORK2Extend.m:68-79 (Heun): heun2 == 0: Ctmp = C - L(C); heun2 == 1:
// C = (C + Ctmp - L(Ctmp)) / 2 with cs = Ctmp, cn = C
__global__ void __launch_bounds__(256) k_enork2_extend_stage(double* __restrict__ out, const double* __restrict__ cs,
                                                             const double* __restrict__ cn,
                                                             const double* __restrict__ vx,
                                                             const double* __restrict__ vy,
                                                             const double* __restrict__ vz, const uint8_t* __restrict__ flags, Dist6c d,
                                                             Dist6c cp,
                                                             Box b, int heun2)
{
#pragma clang fp contract(off)
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    Dists q;
    point_dists(q, d, ind, (flags[ind] & FLAG_SUBCELL) != 0, b);
    const double step = enork2_extend_residual(cs, r, c, p, b, q, cp, ind, vx[ind], vy[ind], vz[ind], 0.3 * min6(q));
    out[ind] = heun2 ? (cn[ind] + cs[ind] - step) / 2 : cs[ind] - step;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define VLS_VARIANT_LAUNCH(kernel, ...)                                                   \
    do {                                                                                  \
        const dim3 block_(64, 4, 1);                                                      \
        const dim3 grid_((g.rows + 63) / 64, (g.cols + 3) / 4, g.pges);                   \
        hipLaunchKernelGGL(kernel, grid_, block_, 0, s, __VA_ARGS__);                     \
        return hipGetLastError();                                                         \
    } while (0)

hipError_t launch_wenork3_re_step(double* step, const double* lsf, const uint8_t* mask, const double* deltat,
                                  const Dist6c& d, const Geom& g, hipStream_t s)
{
    VLS_VARIANT_LAUNCH(k_wenork3_re_step, step, lsf, mask, deltat, d, vbox(g));
}

hipError_t launch_enork2_upwind_normal(double* nx, double* ny, double* nz, const double* lsf, const Dist6c& d,
                                       const Geom& g, hipStream_t s)
{
    VLS_VARIANT_LAUNCH(k_enork2_upwind_normal, nx, ny, nz, lsf, d, vbox(g));
}

hipError_t launch_enork2_boundary_interpolate(const Dist6& cp, const Dist6c& d, const double* lsf, const Geom& g,
                                              hipStream_t s)
{
    VLS_VARIANT_LAUNCH(k_enork2_boundary_interpolate, cp, d, lsf, vbox(g));
}

hipError_t launch_enork2_extend_step(double* step, const double* deltat, const double* lsf, const double* vx,
                                     const double* vy, const double* vz, const Dist6c& d, const Dist6c& cp,
                                     const Geom& g, hipStream_t s)
{
    VLS_VARIANT_LAUNCH(k_enork2_extend_step, step, deltat, lsf, vx, vy, vz, d, cp, vbox(g));
}

hipError_t launch_wenork3_upwind_normal(double* nx, double* ny, double* nz, const double* lsf, const Dist6c& d,
                                        const Geom& g, hipStream_t s)
{
    VLS_VARIANT_LAUNCH(k_wenork3_upwind_normal, nx, ny, nz, lsf, d, vbox(g));
}

hipError_t launch_wenork3_boundary_interpolate(const Dist6& cp, const Dist6c& d, const double* lsf, const Geom& g,
                                               hipStream_t s)
{
    VLS_VARIANT_LAUNCH(k_wenork3_boundary_interpolate, cp, d, lsf, vbox(g));
}

hipError_t launch_wenork3_extend_step(double* step, const double* deltat, const double* lsf, const double* vx,
                                      const double* vy, const double* vz, const Dist6c& d, const Dist6c& cp,
                                      const Geom& g, hipStream_t s)
{
    VLS_VARIANT_LAUNCH(k_wenork3_extend_step, step, deltat, lsf, vx, vy, vz, d, cp, vbox(g));
}

hipError_t launch_enork2_extend_stage(double* out, const double* cs, const double* cn, const double* vx,
                                      const double* vy, const double* vz, const uint8_t* flags, const Dist6c& d,
                                      const Dist6c& cp, const Geom& g, int heun2, hipStream_t s)
{
    VLS_VARIANT_LAUNCH(k_enork2_extend_stage, out, cs, cn, vx, vy, vz, flags, d, cp, vbox(g), heun2);
}

}  // namespace vls
