// Discrete adjoint (reverse mode) and tangent (forward mode) of the FP64
// WENO5 + TVD-RK3 advection step and of the WENO5-RK3 reinitialization with
// its subcell fix, on one grid.
//
// Every fused stage maps its input field fs (and the previous RK level fn)
// to out_i = c_s (fs_i - step_i) + e_s fn_i (c_s, e_s = 1, 0 / 1/4, 3/4 /
// 2/3, 1/3), where step_i depends on fs along the three 7-point lines through
// i (clamped at the box) and on per-point parameters: the velocity for the
// advection; the mask, subcell distances and pseudo-time step of the
// reinitialization.  A Jacobian kernel evaluates the forward formulas on dual
// numbers (vls_dual.hpp), so the WENO weights are differentiated exactly and
// every branch (upwind side, sign-change test, Godunov selection, frozen
// points) is the one the forward pass took, and stores, per point, the 21
// derivatives of out_i along its three lines (rows a*7+k, a = x,y,z, k =
// offset+3) and the parameter derivatives.  The adjoint sweep gathers the
// transposed stencil (deterministic: each output point sums its own
// contributions in a fixed order); the tangent sweep applies it.
//
// The subcell distances of the reinitialization depend on its initial field
// F0 through the root of the quintic through six values (six_distance,
// boundary_location.cu:53-80): d(root)/dy_k = -l_k(root)/p'(root) with l_k the
// Lagrange basis (implicit function theorem); roots clamped to the cell end
// do not move; the quadratic-ENO branch is differentiated on dual numbers.
#include <algorithm>

#include "vls_dual.hpp"
#include "vls_tile.hpp"

namespace vls {

typedef Dual<8> D8;
typedef Dual<9> D9;
typedef Dual<7> D7;

__device__ __forceinline__ double stage_c(int stage) { return stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 2.0 / 3.0); }

// index of the point at `off` from (r,c,p) along axis a, clamped
__device__ __forceinline__ idx_t shifted(const Box& b, int r, int c, int p, int a, int off)
{
    if (a == 0) c = clampi(c + off, b.cols - 1);
    else if (a == 1) r = clampi(r + off, b.rows - 1);
    else p = clampi(p + off, b.pges - 1);
    return lin(b, r, c, p);
}

__device__ __forceinline__ void point_coords(const Box& b, idx_t i, int& r, int& c, int& p) { coords_of(b, i, r, c, p); }

// ---------------------------------------------------------------------------
// advection stage Jacobian: J (21 rows of n), JU (3 rows of n): d out_i / d u_a,i
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) ka_advect_jac(double* __restrict__ J, double* __restrict__ JU,
                                                     const double* __restrict__ fs, const double* __restrict__ vx,
                                                     const double* __restrict__ vy, const double* __restrict__ vz,
                                                     double dt, Box b, Inv inv, int stage)
{
    const idx_t n = (idx_t)b.rows * b.cols * b.pges;
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int r, c, p;
    point_coords(b, i, r, c, p);
    const double cs = stage_c(stage);
    const double vel[3] = {vx[i], vy[i], vz[i]};
#pragma unroll 1
    for (int a = 0; a < 3; ++a) {
        double v[7];
        line7(fs, r, c, p, a, b, v);
        D8 q[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) q[k] = ad::dvar<8>(v[k], k);
        // the term of axis a: coef(u_a) * D_a (upwind_axis), u_a is input 7
        const D8 t = upwind_axis(D8(0.0), ad::dvar<8>(vel[a], 7), q, inv.jd[a]);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            // out = c (fs - dt * sum_a t_a) + e fn
            double j = -cs * dt * t.d[k];
            if (a == 0 && k == 3) j += cs;
            J[(size_t)(a * 7 + k) * n + i] = j;
        }
        JU[(size_t)a * n + i] = -cs * dt * t.d[7];
    }
}

// ---------------------------------------------------------------------------
// reinitialization stage Jacobian (re_step.cu:283-393 as the fused stage
// kernels evaluate it).  Points without subcell distances: spacing
// distances and the bulk time step, no parameter derivatives.  Listed points:
// distances and time step from the list, and JP (7 rows of cap, list order):
// d out / d (xpr, xpl, ypf, ypb, zpu, zpd, dt).
// ---------------------------------------------------------------------------
// axis pair (fore, back) on duals of the 7 line values (inputs 0..6) and
// r1, l1 (inputs 7, 8), with reinit_axis's branch
__device__ __forceinline__ void reinit_axis_dual(D9& fo, D9& bk, const double v[7], double r1, double l1, double ds,
                                                 JPd jp)
{
    D9 q[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = ad::dvar<9>(v[k], k);
    if (v[2] * v[3] < 0 || v[3] * v[4] < 0) {
        const PairT<D9> pr = weno_nonuniform_pair(q, ad::dvar<9>(r1, 7), ad::dvar<9>(l1, 8), ds, D9(0.0), D9(0.0));
        fo = pr.fore;
        bk = pr.back;
    } else {
        weno5_pair(fo, bk, q, jp);
    }
}

// one point: fills its 21 J entries (and, when jp is not null, its 7
// parameter derivatives); dt < 0 is a frozen point (step 0)
__device__ __forceinline__ void reinit_point_jac(double* __restrict__ J, double* jp, idx_t n, idx_t i,
                                                 const double* __restrict__ fs, const Box& b, const Inv& inv,
                                                 const double r1[3], const double l1[3], double dt, bool inside,
                                                 int stage)
{
    int r, c, p;
    point_coords(b, i, r, c, p);
    const double cs = stage_c(stage);
    const double sp[3] = {b.dx, b.dy, b.dz};
    if (!(dt >= 0)) {
        for (int e = 0; e < 21; ++e) J[(size_t)e * n + i] = (e == 3) ? cs : 0.0;
        if (jp)
            for (int e = 0; e < 7; ++e) jp[e] = 0.0;
        return;
    }
    D9 fo[3], bk[3];
    double v[3][7];
#pragma unroll 1
    for (int a = 0; a < 3; ++a) {
        line7(fs, r, c, p, a, b, v[a]);
        reinit_axis_dual(fo[a], bk[a], v[a], r1[a], l1[a], sp[a], inv.jd[a]);
    }
    // Godunov step on duals of its 7 arguments (fore/back per axis, dt)
    const D7 g = godunov_step(inside, ad::dvar<7>(fo[0].v, 0), ad::dvar<7>(bk[0].v, 1), ad::dvar<7>(fo[1].v, 2),
                              ad::dvar<7>(bk[1].v, 3), ad::dvar<7>(fo[2].v, 4), ad::dvar<7>(bk[2].v, 5),
                              ad::dvar<7>(dt, 6));
#pragma unroll 1
    for (int a = 0; a < 3; ++a) {
        for (int k = 0; k < 7; ++k) {
            double j = -cs * (g.d[2 * a] * fo[a].d[k] + g.d[2 * a + 1] * bk[a].d[k]);
            if (a == 0 && k == 3) j += cs;
            J[(size_t)(a * 7 + k) * n + i] = j;
        }
        if (jp) {
            jp[2 * a] = -cs * (g.d[2 * a] * fo[a].d[7] + g.d[2 * a + 1] * bk[a].d[7]);
            jp[2 * a + 1] = -cs * (g.d[2 * a] * fo[a].d[8] + g.d[2 * a + 1] * bk[a].d[8]);
        }
    }
    if (jp) jp[6] = -cs * g.d[6];
}

__global__ void __launch_bounds__(256) ka_reinit_jac_bulk(double* __restrict__ J, const double* __restrict__ fs,
                                                          const uint8_t* __restrict__ flags, Box b, Inv inv,
                                                          int stage)
{
    const idx_t n = (idx_t)b.rows * b.cols * b.pges;
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t fl = flags[i];
    if (fl & FLAG_SUBCELL) return;
    const double r1[3] = {b.dx, b.dy, b.dz};
    reinit_point_jac(J, nullptr, n, i, fs, b, inv, r1, r1, inv.dt_bulk, (fl & FLAG_MASK) != 0, stage);
}

__global__ void __launch_bounds__(256) ka_reinit_jac_list(double* __restrict__ J, double* __restrict__ JP,
                                                          const double* __restrict__ fs,
                                                          const uint8_t* __restrict__ flags, IList il, Box b, Inv inv,
                                                          int stage)
{
    const idx_t n = (idx_t)b.rows * b.cols * b.pges;
    const int cnt = *il.count;
    for (int j = (int)(blockIdx.x * blockDim.x + threadIdx.x); j < cnt; j += (int)(gridDim.x * blockDim.x)) {
        const idx_t i = il.idx[j];
        double r1[3], l1[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            r1[a] = il.dist[(size_t)(2 * a) * il.cap + j];
            l1[a] = il.dist[(size_t)(2 * a + 1) * il.cap + j];
        }
        double jp[7];
        reinit_point_jac(J, jp, n, i, fs, b, inv, r1, l1, il.dt[j], (flags[i] & FLAG_MASK) != 0, stage);
#pragma unroll
        for (int e = 0; e < 7; ++e) JP[(size_t)e * il.cap + j] = jp[e];
    }
}

// ---------------------------------------------------------------------------
// sweeps
// ---------------------------------------------------------------------------
// adjoint: lam_in[j] = beta*lam_in[j] + sum over (i, a, k) with
// clamp(i + (k-3) e_a) = j of J[a*7+k][i] lam[i]
__global__ void __launch_bounds__(256) ka_transpose(double* __restrict__ lam_in, const double* __restrict__ J,
                                                    const double* __restrict__ lam, Box b, double beta)
{
    const idx_t n = (idx_t)b.rows * b.cols * b.pges;
    const idx_t j = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int r, c, p;
    point_coords(b, j, r, c, p);
    double acc = 0.0;
#pragma unroll 1
    for (int a = 0; a < 3; ++a) {
        const int ja = (a == 0) ? c : (a == 1 ? r : p);
        const int Na = (a == 0) ? b.cols : (a == 1 ? b.rows : b.pges);
#pragma unroll 1
        for (int k = 0; k < 7; ++k) {
            const int o = k - 3;
            int lo, hi;  // the i_a with clamp(i_a + o) == ja
            if (ja > 0 && ja < Na - 1) {
                lo = hi = ja - o;
            } else if (ja == 0) {
                lo = 0;
                hi = -o;
            } else {
                lo = Na - 1 - o;
                hi = Na - 1;
            }
            lo = lo < 0 ? 0 : lo;
            hi = hi > Na - 1 ? Na - 1 : hi;
            for (int ia = lo; ia <= hi; ++ia) {
                const idx_t i = shifted(b, r, c, p, a, ia - ja);
                acc += J[(size_t)(a * 7 + k) * n + i] * lam[i];
            }
        }
    }
    lam_in[j] = (beta == 0.0 ? 0.0 : beta * lam_in[j]) + acc;
}

// tangent: out[i] = sum_{a,k} J[a*7+k][i] x[clamp(i + (k-3) e_a)]
__global__ void __launch_bounds__(256) ka_apply(double* __restrict__ out, const double* __restrict__ J,
                                                const double* __restrict__ x, Box b)
{
    const idx_t n = (idx_t)b.rows * b.cols * b.pges;
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int r, c, p;
    point_coords(b, i, r, c, p);
    double acc = 0.0;
#pragma unroll 1
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 7; ++k) acc += J[(size_t)(a * 7 + k) * n + i] * x[shifted(b, r, c, p, a, k - 3)];
    out[i] = acc;
}

__global__ void ka_fill(double* __restrict__ y, double v, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

// y = alpha*x + beta*y (beta == 0 overwrites)
__global__ void ka_axpby(double* __restrict__ y, const double* __restrict__ x, double alpha, double beta, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = alpha * x[i] + (beta == 0.0 ? 0.0 : beta * y[i]);
}

// g_a[i] += JU[a][i] * lam[i]   (velocity gradient)
__global__ void ka_velocity_grad(double* __restrict__ gx, double* __restrict__ gy, double* __restrict__ gz,
                                 const double* __restrict__ JU, const double* __restrict__ lam, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double l = lam[i];
    gx[i] += JU[i] * l;
    gy[i] += JU[(size_t)n + i] * l;
    gz[i] += JU[(size_t)2 * n + i] * l;
}

// out[i] += sum_a JU[a][i] * udot_a[i]   (velocity tangent)
__global__ void ka_velocity_tangent(double* __restrict__ out, const double* __restrict__ JU,
                                    const double* __restrict__ ux, const double* __restrict__ uy,
                                    const double* __restrict__ uz, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] += JU[i] * ux[i] + JU[(size_t)n + i] * uy[i] + JU[(size_t)2 * n + i] * uz[i];
}

// reinitialization parameters, list order (7 rows of cap):
// adjoint  lamP[e][j] += JP[e][j] * lam[idx[j]];  tangent out[idx[j]] += sum_e JP[e][j] Pdot[e][j]
__global__ void ka_param_grad(double* __restrict__ lamP, const double* __restrict__ JP, const double* __restrict__ lam,
                              IList il)
{
    const int cnt = *il.count;
    for (int j = (int)(blockIdx.x * blockDim.x + threadIdx.x); j < cnt; j += (int)(gridDim.x * blockDim.x)) {
        const double l = lam[il.idx[j]];
        for (int e = 0; e < 7; ++e) lamP[(size_t)e * il.cap + j] += JP[(size_t)e * il.cap + j] * l;
    }
}

__global__ void ka_param_tangent(double* __restrict__ out, const double* __restrict__ JP,
                                 const double* __restrict__ Pdot, IList il)
{
    const int cnt = *il.count;
    for (int j = (int)(blockIdx.x * blockDim.x + threadIdx.x); j < cnt; j += (int)(gridDim.x * blockDim.x)) {
        double s = 0.0;
        for (int e = 0; e < 7; ++e) s += JP[(size_t)e * il.cap + j] * Pdot[(size_t)e * il.cap + j];
        out[il.idx[j]] += s;
    }
}

// ---------------------------------------------------------------------------
// subcell distances as functions of F0
// ---------------------------------------------------------------------------
// eno_quadratic_distance (boundary_location.cu:6-20) on duals
template <int N>
__device__ __forceinline__ Dual<N> eno_quadratic_distance_d(Dual<N> v1, Dual<N> v0, Dual<N> v2, Dual<N> v3, double ds)
{
#pragma clang fp contract(off)
    const double eps = 10e-10;
    const Dual<N> p2l = v2 - 2.0 * v0 + v1;
    const Dual<N> p2r = v3 - 2.0 * v2 + v0;
    // min_mod on values, derivative of the chosen argument
    Dual<N> p2;
    if (p2l.v * p2r.v < 0) p2 = Dual<N>(0.0);
    else p2 = (fabs(p2l.v) < fabs(p2r.v)) ? p2l : p2r;
    if (p2.v > eps || p2.v < -eps) {
        const Dual<N> disc = sq(0.5 * p2 - v0 - v2) - 4. * v0 * v2;
        return ds * (0.5 + (v0 - v2 - ksign(v0.v - v2.v) * sqrt(disc)) / p2);
    }
    return ds * v0 / (v0 - v2);
}

// d(distance)/d(y_0..y_5) of six_distance(y_0..y_5, s); returns false when the
// distance does not move (no crossing or a root clamped to the cell end)
__device__ bool six_distance_grad(const double y[6], double s, double g[6])
{
    for (int k = 0; k < 6; ++k) g[k] = 0.0;
    if (y[1] * y[2] < 0 || y[3] * y[4] < 0) {
        Dual<4> q[4];
        for (int k = 0; k < 4; ++k) q[k] = ad::dvar<4>(y[k + 1], k);
        const Dual<4> d = eno_quadratic_distance_d(q[0], q[1], q[2], q[3], s);
        for (int k = 0; k < 4; ++k) g[k + 1] = d.d[k];
        return true;
    }
    const double xx[6] = {0.0, s, 2 * s, 3 * s, 4 * s, 5 * s};
    double coef[6];
    poly5_interp(xx, y, coef);
    const double x = zeroin_poly5(coef, 2 * s, 3 * s, 3e-16);
    if (!(x > 2 * s && x < 3 * s)) return false;
    // p'(x) from the monomial coefficients, Lagrange basis at x
    double dp = 0.0, pw = 1.0;
    for (int m = 1; m <= 5; ++m) {
        dp += m * coef[m] * pw;
        pw *= x;
    }
    if (dp == 0.0) return false;
    for (int k = 0; k < 6; ++k) {
        double l = 1.0;
        for (int m = 0; m < 6; ++m)
            if (m != k) l *= (x - xx[m]) / (xx[k] - xx[m]);
        g[k] = -l / dp;
    }
    return true;
}

// the six distances of a listed point i as functions of F0: for each of the
// 12 (axis, side) pairs, the six F0 indices and derivatives
__device__ void dist_grads(const double* __restrict__ F0, const Box& b, idx_t i, int a, int side, idx_t idx[6],
                           double g[6], bool& moves)
{
    int r, c, p;
    point_coords(b, i, r, c, p);
    double v[7];
    line7(F0, r, c, p, a, b, v);
    const double sp[3] = {b.dx, b.dy, b.dz};
    moves = false;
    if (side == 0) {  // r-side: crossing (i, i+1), stencil i-2..i+3
        const idx_t nb = shifted(b, r, c, p, a, 1);
        if (!(v[3] * F0[nb] < 0)) return;
        for (int k = 0; k < 6; ++k) idx[k] = shifted(b, r, c, p, a, k - 2);
        const double y[6] = {v[1], v[2], v[3], v[4], v[5], v[6]};
        moves = six_distance_grad(y, sp[a], g);
    } else {  // l-side: ds - distance of the crossing (i-1, i), stencil i-3..i+2
        const idx_t nb = shifted(b, r, c, p, a, -1);
        if (!(F0[nb] * v[3] < 0)) return;
        for (int k = 0; k < 6; ++k) idx[k] = shifted(b, r, c, p, a, k - 3);
        const double y[6] = {v[0], v[1], v[2], v[3], v[4], v[5]};
        moves = six_distance_grad(y, sp[a], g);
        for (int k = 0; k < 6; ++k) g[k] = -g[k];
    }
}

// which of the six distances the time step 0.3*min_dist follows (min6's order)
__device__ __forceinline__ int argmin6(const double d[6])
{
    // min6 = dmin(dmin(xr, xl), dmin(dmin(yr, yl), dmin(zr, zl))) with d = {xr, xl, yr, yl, zr, zl}
    auto pick = [](int a, int b2, const double* e) { return (e[a] < e[b2]) ? a : b2; };
    const int x = pick(0, 1, d), y = pick(2, 3, d), z = pick(4, 5, d);
    const int yz = pick(y, z, d);
    return pick(x, yz, d);
}

// adjoint of the distances: lam_F0[stencil] += d(dist)/dF0 * lam_dist, where
// lam_dist[2a+side] = lamP[2a+side] + (argmin == 2a+side) * 0.3 * lamP[6]
__global__ void __launch_bounds__(256) ka_dist_adjoint(double* __restrict__ lamF0, const double* __restrict__ F0,
                                                       const double* __restrict__ lamP, IList il, Box b)
{
    const int cnt = *il.count;
    for (int j = (int)(blockIdx.x * blockDim.x + threadIdx.x); j < cnt; j += (int)(gridDim.x * blockDim.x)) {
        const idx_t i = il.idx[j];
        double d[6];
        for (int e = 0; e < 6; ++e) d[e] = il.dist[(size_t)e * il.cap + j];
        const bool frozen_pt = !(il.dt[j] >= 0);
        const int am = argmin6(d);
        for (int e = 0; e < 6; ++e) {
            double lam = lamP[(size_t)e * il.cap + j];
            if (!frozen_pt && e == am) lam += 0.3 * lamP[(size_t)6 * il.cap + j];
            if (lam == 0.0) continue;
            idx_t idx[6];
            double g[6];
            bool moves;
            dist_grads(F0, b, i, e >> 1, e & 1, idx, g, moves);
            if (!moves) continue;
            for (int k = 0; k < 6; ++k) atomicAdd(&lamF0[idx[k]], g[k] * lam);
        }
    }
}

// tangent of the parameters: Pdot[e][j] for the distances and the time step
__global__ void __launch_bounds__(256) ka_dist_tangent(double* __restrict__ Pdot, const double* __restrict__ F0,
                                                       const double* __restrict__ F0dot, IList il, Box b)
{
    const int cnt = *il.count;
    for (int j = (int)(blockIdx.x * blockDim.x + threadIdx.x); j < cnt; j += (int)(gridDim.x * blockDim.x)) {
        const idx_t i = il.idx[j];
        double d[6], dd[6];
        for (int e = 0; e < 6; ++e) {
            d[e] = il.dist[(size_t)e * il.cap + j];
            idx_t idx[6];
            double g[6];
            bool moves;
            dist_grads(F0, b, i, e >> 1, e & 1, idx, g, moves);
            double s = 0.0;
            if (moves)
                for (int k = 0; k < 6; ++k) s += g[k] * F0dot[idx[k]];
            dd[e] = s;
            Pdot[(size_t)e * il.cap + j] = s;
        }
        Pdot[(size_t)6 * il.cap + j] = (il.dt[j] >= 0) ? 0.3 * dd[argmin6(d)] : 0.0;
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }
static inline Box abox(const Geom& g) { return Box{g.rows, g.cols, g.pges, g.dx, g.dy, g.dz}; }
static inline Inv ainv(const Geom& g)
{
    const double m = std::min(std::min(g.dx, g.dx), std::min(std::min(g.dy, g.dy), std::min(g.dz, g.dz)));
    Inv v{{1.0 / g.dx, 1.0 / g.dy, 1.0 / g.dz}, 0.3 * m, {}, {}};
    set_jp(v.jd, v.jf, g.dx, g.dy, g.dz);
    return v;
}
static inline IList alist(const ListView& lv) { return IList{lv.idx, lv.count, lv.dist, lv.dt, lv.cap}; }

hipError_t launch_adj_advect_jac(double* J, double* JU, const double* fs, const double* vx, const double* vy,
                                 const double* vz, double dt, const Geom& g, int stage, hipStream_t s)
{
    ka_advect_jac<<<grid1(g.n()), 256, 0, s>>>(J, JU, fs, vx, vy, vz, dt, abox(g), ainv(g), stage);
    return hipGetLastError();
}

hipError_t launch_adj_reinit_jac(double* J, double* JP, const double* fs, const uint8_t* flags, const ListView& lv,
                                 const Geom& g, int stage, hipStream_t s)
{
    ka_reinit_jac_bulk<<<grid1(g.n()), 256, 0, s>>>(J, fs, flags, abox(g), ainv(g), stage);
    ka_reinit_jac_list<<<256, 256, 0, s>>>(J, JP, fs, flags, alist(lv), abox(g), ainv(g), stage);
    return hipGetLastError();
}

hipError_t launch_adj_transpose(double* lam_in, const double* J, const double* lam, const Geom& g, double beta,
                                hipStream_t s)
{
    ka_transpose<<<grid1(g.n()), 256, 0, s>>>(lam_in, J, lam, abox(g), beta);
    return hipGetLastError();
}

hipError_t launch_adj_apply(double* out, const double* J, const double* x, const Geom& g, hipStream_t s)
{
    ka_apply<<<grid1(g.n()), 256, 0, s>>>(out, J, x, abox(g));
    return hipGetLastError();
}

hipError_t launch_adj_axpby(double* y, const double* x, double alpha, double beta, int64_t n, hipStream_t s)
{
    ka_axpby<<<grid1(n), 256, 0, s>>>(y, x, alpha, beta, (idx_t)n);
    return hipGetLastError();
}

hipError_t launch_adj_fill(double* y, double v, int64_t n, hipStream_t s)
{
    ka_fill<<<grid1(n), 256, 0, s>>>(y, v, (idx_t)n);
    return hipGetLastError();
}

hipError_t launch_adj_velocity_grad(double* gx, double* gy, double* gz, const double* JU, const double* lam,
                                    int64_t n, hipStream_t s)
{
    ka_velocity_grad<<<grid1(n), 256, 0, s>>>(gx, gy, gz, JU, lam, (idx_t)n);
    return hipGetLastError();
}

hipError_t launch_adj_velocity_tangent(double* out, const double* JU, const double* ux, const double* uy,
                                       const double* uz, int64_t n, hipStream_t s)
{
    ka_velocity_tangent<<<grid1(n), 256, 0, s>>>(out, JU, ux, uy, uz, (idx_t)n);
    return hipGetLastError();
}

hipError_t launch_adj_param_grad(double* lamP, const double* JP, const double* lam, const ListView& lv, hipStream_t s)
{
    ka_param_grad<<<256, 256, 0, s>>>(lamP, JP, lam, alist(lv));
    return hipGetLastError();
}

hipError_t launch_adj_param_tangent(double* out, const double* JP, const double* Pdot, const ListView& lv,
                                    hipStream_t s)
{
    ka_param_tangent<<<256, 256, 0, s>>>(out, JP, Pdot, alist(lv));
    return hipGetLastError();
}

hipError_t launch_adj_dist(double* lamF0, const double* F0, const double* lamP, const ListView& lv, const Geom& g,
                           hipStream_t s)
{
    ka_dist_adjoint<<<256, 256, 0, s>>>(lamF0, F0, lamP, alist(lv), abox(g));
    return hipGetLastError();
}

hipError_t launch_tan_dist(double* Pdot, const double* F0, const double* F0dot, const ListView& lv, const Geom& g,
                           hipStream_t s)
{
    ka_dist_tangent<<<256, 256, 0, s>>>(Pdot, F0, F0dot, alist(lv), abox(g));
    return hipGetLastError();
}

}  // namespace vls
