// HIP kernels of the level-set hot path for gfx950.
//
// Layout: double[rows*cols*pges], ind = p*rows*cols + c*rows + r (rows = y is
// the contiguous axis).  A block is 64 rows x 4 cols of one z-page: every
// wave64 reads 64 consecutive doubles (one 512-byte segment) per stencil tap
// along x and z, and the +-3 y-taps of a wave fall in the same or the two
// neighbouring segments.  All neighbour indices are clamped to the box.
//
// Two families:
//   * drop-in kernels with the exact argument meaning of the reference's
//     __global__ entry points (re_step, extend_step, ...);
//   * fused stage kernels that evaluate one operator and its Shu-Osher RK
//     combination in one pass (the reference runs the combination as separate
//     MATLAB gpuArray kernels), read the six subcell distances only where a
//     per-point flag says they differ from the grid spacing, and derive the
//     pseudo-time step 0.3*min_dist from them.
#include <algorithm>

#include "vls_tile.hpp"

namespace vls {


// ---------------------------------------------------------------------------
// setDistance: fill, crossing distances, min_dist + flags
// ---------------------------------------------------------------------------
// boundary_location (order 6, 1_3_WENO5RK3_Reinitialization/boundary_location.cu:84-132),
// boundary_correction (order 2, 1_0_ENORK2_Reinitialization/boundary_correction.cu:103-145)
// and the cubic boundary_correction (order 3,
// 1_1_WENORK3_Reinitialization/weno_rk3_reinitialization.cu:251-292).
// One writer per element: r-side at ind, l-side at the +1 neighbour.
template <int ORDER, typename T>
__global__ void __launch_bounds__(256) k_boundary_location(Dist6 d, const T* __restrict__ lsf, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double f0 = (double)lsf[ind];
    const double sp[3] = {b.dx, b.dy, b.dz};
#pragma unroll 1
    for (int axis = 0; axis < 3; ++axis) {
        int rr = r, cc = c, pp = p;
        if (axis == 0) cc = min(c + 1, b.cols - 1);
        else if (axis == 1) rr = min(r + 1, b.rows - 1);
        else pp = min(p + 1, b.pges - 1);
        const idx_t nb = lin(b, rr, cc, pp);
        const double f2 = (double)lsf[nb];
        if (!(f0 * f2 < 0)) continue;
        double v[7];
        line7(lsf, r, c, p, axis, b, v);
        const double ds = sp[axis];
        const double dist = (ORDER == 2)   ? eno_quadratic_distance(v[2], f0, f2, v[5], ds)
                            : (ORDER == 3) ? cubic_distance(v[2], f0, f2, v[5], ds)
                                           : six_distance(v[1], v[2], f0, f2, v[5], v[6], ds);
        d.r[axis][ind] = dist;
        d.l[axis][nb] = ds - dist;
    }
}


// setDistance (setDistance.m:3-21) in two passes.  The first visits every
// point: it tests the crossings with the six neighbours (F(i) F(i+-1) < 0 on
// an axis, the locators' own tests), writes the flags without FLAG_SUBCELL
// (FLAG_MASK where mask_src < 0) and appends the points with a crossing to
// `cand`.  The second visits those points only: it locates their crossings,
// and the points where a distance then differs from the spacing get
// FLAG_SUBCELL and enter the interface list, in list order, with their six
// distances and pseudo-time step (the interface kernels read them
// coalesced).  The dense arrays (six distances and min_dist of every point,
// the MATLAB object's properties) are written only where asked for (d.r[0]
// set): spacings by the first pass, located distances by the second.
//
// the six neighbour values of a point (clamped at the box faces)
template <typename T>
__device__ __forceinline__ void neighbours6(const T* __restrict__ lsf, const Box& b, idx_t ind, int r, int c, int p,
                                            double up[3], double dn[3])
{
    const idx_t rows = (idx_t)b.rows, plane = rows * b.cols;
    up[0] = (double)lsf[ind + (idx_t)(clampi(c + 1, b.cols - 1) - c) * rows];
    dn[0] = (double)lsf[ind + (idx_t)(clampi(c - 1, b.cols - 1) - c) * rows];
    up[1] = (double)lsf[ind + (idx_t)(clampi(r + 1, b.rows - 1) - r)];
    dn[1] = (double)lsf[ind + (idx_t)(clampi(r - 1, b.rows - 1) - r)];
    up[2] = (double)lsf[ind + (idx_t)(clampi(p + 1, b.pges - 1) - p) * plane];
    dn[2] = (double)lsf[ind + (idx_t)(clampi(p - 1, b.pges - 1) - p) * plane];
}

template <typename T>
__global__ void __launch_bounds__(256) k_crossings(Dist6 d, double* __restrict__ min_dist,
                                                   uint8_t* __restrict__ flags, const T* __restrict__ mask_src,
                                                   const T* __restrict__ lsf, Box b, int* __restrict__ cand,
                                                   int* __restrict__ cand_count)
{
    int r, c, p;
    const bool in = point_of_thread(b, r, c, p);
    bool cross = false;
    idx_t ind = 0;
    if (in) {
        ind = lin(b, r, c, p);
        const double f0 = (double)lsf[ind];
        double up[3], dn[3];
        neighbours6(lsf, b, ind, r, c, p, up, dn);
#pragma unroll
        for (int a = 0; a < 3; ++a) cross |= (f0 * up[a] < 0) | (dn[a] * f0 < 0);
        if (flags) flags[ind] = (mask_src && mask_src[ind] < 0) ? (uint8_t)FLAG_MASK : (uint8_t)0;
        if (d.r[0]) {
            d.r[0][ind] = b.dx; d.l[0][ind] = b.dx;
            d.r[1][ind] = b.dy; d.l[1][ind] = b.dy;
            d.r[2][ind] = b.dz; d.l[2][ind] = b.dz;
            if (min_dist) min_dist[ind] = dmin(dmin(b.dx, b.dx), dmin(dmin(b.dy, b.dy), dmin(b.dz, b.dz)));
        }
    }
    wave_append(cross, (int)ind, cand, cand_count);
}

// The first pass on the z-marching walk of the stage kernels: the x- and
// y-neighbours come from the staged tile, the z-neighbours from the register
// window, so each value of F is read from memory about once.  mask: FLAG_MASK
// where F < 0 (the only mask source the setDistance calls use is F itself).
template <int R, typename T>
__global__ void __launch_bounds__(256) k_crossings_z(Dist6 d, double* __restrict__ min_dist,
                                                     uint8_t* __restrict__ flags, int mask,
                                                     const T* __restrict__ lsf, Box b, int depth,
                                                     int* __restrict__ cand, int* __restrict__ cand_count)
{
    __shared__ T lds[2][Shape<R>::N];
    __shared__ int cbuf[1024], ccount, cbase;
    BlockList<1024> bl{cbuf, &ccount, &cbase};
    bl.init();
    int r0, c0, p0, p1;
    zmarch_share<R>(b, depth, (int)blockIdx.x, (int)gridDim.x, r0, c0, p0, p1);
    const idx_t plane = (idx_t)b.rows * b.cols;
    const double md = dmin(dmin(b.dx, b.dx), dmin(dmin(b.dy, b.dy), dmin(b.dz, b.dz)));
    zmarch<R, T, false, false>(lds[0], lds[1], b, r0, c0, p0, p1, lsf, nullptr, nullptr, nullptr, nullptr, nullptr,
                               0, [&](const T* l, const ZCol<R>& zc, const T w[7], const ZStep<R, T, false>&, int p) {
                                   const double f0 = (double)w[3];
                                   const double xu = (double)l[zc.base + Shape<R>::LR],
                                                xd = (double)l[zc.base - Shape<R>::LR];
                                   const double yu = (double)l[zc.base + 1], yd = (double)l[zc.base - 1];
                                   const double zu = (double)w[4], zd = (double)w[2];
                                   const bool cross = zc.valid && ((f0 * xu < 0) | (xd * f0 < 0) | (f0 * yu < 0) |
                                                                   (yd * f0 < 0) | (f0 * zu < 0) | (zd * f0 < 0));
                                   const idx_t ind = (idx_t)p * plane + zc.in;
                                   if (zc.valid) {
                                       if (flags) flags[ind] = (mask && w[3] < 0) ? (uint8_t)FLAG_MASK : (uint8_t)0;
                                       if (d.r[0]) {
                                           d.r[0][ind] = b.dx; d.l[0][ind] = b.dx;
                                           d.r[1][ind] = b.dy; d.l[1][ind] = b.dy;
                                           d.r[2][ind] = b.dz; d.l[2][ind] = b.dz;
                                           if (min_dist) min_dist[ind] = md;
                                       }
                                   }
                                   bl.push(cross, (int)ind, cand, cand_count);
                               });
    bl.flush(cand, cand_count);
}

// the second pass, one lane per candidate.  The r-side distance on an axis is
// the locator at the point (crossing with the +1 neighbour); the l-side one
// is ds minus the locator of the -1 neighbour's crossing, evaluated here
// from that neighbour's stencil (the point's own line shifted by one,
// clamped the same way), so it is the value k_boundary_location stores from
// the neighbour's thread.  FLAG_SUBCELL marks a distance that differs from
// the spacing: below it, or, for the quadratic ENO locator, which is not
// clamped, a root within rounding of the far node that lands just above it,
// where the ENO2 schemes treat it as a crossing.
template <int ORDER, typename T>
__global__ void __launch_bounds__(256) k_crossing_distances(Dist6 d, double* __restrict__ min_dist,
                                                            uint8_t* __restrict__ flags, const T* __restrict__ lsf,
                                                            Box b, const int* __restrict__ cand,
                                                            const int* __restrict__ cand_count, ListView lv)
{
    const int total = *cand_count;
    const double sp[3] = {b.dx, b.dy, b.dz};
    // whole waves per sweep, so that the list append sees every lane
    for (int j0 = (int)(blockIdx.x * 256); j0 < total; j0 += (int)(gridDim.x * 256)) {
        const int j = j0 + (int)threadIdx.x;
        bool sub = false;
        idx_t ind = 0;
        Dists q;
        if (j < total) {
            ind = (idx_t)cand[j];
            int r, c, p;
            coords_of(b, ind, r, c, p);
            const idx_t plane = (idx_t)b.rows * b.cols;
            const double f0 = (double)lsf[ind];
#pragma unroll 1
            for (int a = 0; a < 3; ++a) {
                const double ds = sp[a];
                q.r[a] = ds;
                q.l[a] = ds;
                const int pos = a == 0 ? c : (a == 1 ? r : p);
                const int last = a == 0 ? b.cols - 1 : (a == 1 ? b.rows - 1 : b.pges - 1);
                const idx_t stride = a == 0 ? (idx_t)b.rows : (a == 1 ? (idx_t)1 : plane);
                const idx_t base = ind - (idx_t)pos * stride;
                double v[8];  // the line at offsets -4..+3, clamped
#pragma unroll
                for (int s = -4; s <= 3; ++s) v[s + 4] = (double)lsf[base + (idx_t)clampi(pos + s, last) * stride];
                const double fu = v[5], fd = v[3];
                // the locator at the point: its line is v[1..7], its f2 = fu
                if (f0 * fu < 0)
                    q.r[a] = (ORDER == 2)   ? eno_quadratic_distance(v[3], f0, fu, v[6], ds)
                             : (ORDER == 3) ? cubic_distance(v[3], f0, fu, v[6], ds)
                                            : six_distance(v[2], v[3], f0, fu, v[6], v[7], ds);
                // the -1 neighbour's locator: its line is v[0..6], its f0 = fd, its f2 = f0
                if (fd * f0 < 0)
                    q.l[a] = ds - ((ORDER == 2)   ? eno_quadratic_distance(v[2], fd, f0, v[5], ds)
                                   : (ORDER == 3) ? cubic_distance(v[2], fd, f0, v[5], ds)
                                                  : six_distance(v[1], v[2], fd, f0, v[5], v[6], ds));
            }
            sub = q.r[0] != b.dx || q.l[0] != b.dx || q.r[1] != b.dy || q.l[1] != b.dy || q.r[2] != b.dz ||
                  q.l[2] != b.dz;
            if (sub) {
                if (flags) flags[ind] |= (uint8_t)FLAG_SUBCELL;
                if (d.r[0]) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        d.r[a][ind] = q.r[a];
                        d.l[a][ind] = q.l[a];
                    }
                    if (min_dist) min_dist[ind] = min6(q);
                }
            }
        }
        if (!lv.idx) continue;
        const int slot = wave_append_slot(sub, lv.count);
        if (slot < 0) continue;
        lv.idx[slot] = (int)ind;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lv.dist[(size_t)(2 * a) * lv.cap + slot] = q.r[a];
            lv.dist[(size_t)(2 * a + 1) * lv.cap + slot] = q.l[a];
        }
        lv.dt[slot] = frozen(q, b.dx) ? -1.0 : 0.3 * min6(q);
    }
}

// the dense arrays of a lazily located distance set: spacings everywhere,
// then the listed points' distances (every point with a distance that
// differs from the spacing is listed, with all six)
__global__ void k_fill_dense(Dist6 d, double* __restrict__ min_dist, Box b, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d.r[0][i] = b.dx; d.l[0][i] = b.dx;
    d.r[1][i] = b.dy; d.l[1][i] = b.dy;
    d.r[2][i] = b.dz; d.l[2][i] = b.dz;
    min_dist[i] = dmin(dmin(b.dx, b.dx), dmin(dmin(b.dy, b.dy), dmin(b.dz, b.dz)));
}

__global__ void __launch_bounds__(256) k_scatter_list(Dist6 d, double* __restrict__ min_dist, IList il)
{
    const int cnt = *il.count;
    for (int j = (int)(blockIdx.x * 256 + threadIdx.x); j < cnt; j += (int)(gridDim.x * 256)) {
        const idx_t ind = (idx_t)il.idx[j];
        Dists q;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            q.r[a] = il.dist[(size_t)(2 * a) * il.cap + j];
            q.l[a] = il.dist[(size_t)(2 * a + 1) * il.cap + j];
            d.r[a][ind] = q.r[a];
            d.l[a][ind] = q.l[a];
        }
        min_dist[ind] = min6(q);
    }
}

__global__ void k_zero_pair(int* __restrict__ a, int* __restrict__ b)
{
    if (threadIdx.x == 0) {
        *a = 0;
        if (b) *b = 0;
    }
}

__global__ void k_zero_ints(int* __restrict__ a, int n)
{
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0;
}

// mask bit only (reinit with distances from an earlier setDistance)
template <typename T>
__global__ void k_set_mask_bit(uint8_t* __restrict__ flags, const T* __restrict__ F, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t f = flags[i] & (uint8_t)~FLAG_MASK;
    if (F[i] < 0) f |= FLAG_MASK;
    flags[i] = f;
}

// ---------------------------------------------------------------------------
// WENO5 reinitialization residual (re_step.cu:283-393)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double reinit_residual(const double* __restrict__ f, int r, int c, int p, const Box& b,
                                                  const Dists& q, bool inside, double dt)
{
    const double sp[3] = {b.dx, b.dy, b.dz};
    double fo[3], bk[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7];
        line7(f, r, c, p, a, b, v);
        reinit_axis(fo[a], bk[a], v, q.r[a], q.l[a], sp[a], jpd(sp[a]));
    }
    return godunov_step(inside, fo[0], bk[0], fo[1], bk[1], fo[2], bk[2], dt);
}

// drop-in re_step
__global__ void __launch_bounds__(256) k_re_step(double* __restrict__ step, const double* __restrict__ lsf,
                                                 const uint8_t* __restrict__ mask,
                                                 const double* __restrict__ deltat, Dist6c d, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    Dists q;
#pragma unroll
    for (int a = 0; a < 3; ++a) { q.r[a] = d.r[a][ind]; q.l[a] = d.l[a][ind]; }
    if (frozen(q, b.dx)) { step[ind] = 0; return; }
    step[ind] = reinit_residual(lsf, r, c, p, b, q, mask[ind] != 0, deltat[ind]);
}

// ---------------------------------------------------------------------------
// split stage kernels.  Bulk kernels cover the box in 64x4 (rows x cols)
// tiles of one page and take every point without subcell distances through
// uniform WENO5 only (few registers, high occupancy); interface kernels
// stride over a list of points and evaluate the full residual.
// ---------------------------------------------------------------------------



// Fused WENO5-RK3 reinitialization stage, launch 1 of 2: the first niface
// blocks walk the interface list with three lanes per point (the full
// residual with subcell distances), the others z-march over every point
// without subcell distances.  The list blocks are dispatched first, so their
// long dependent chains (the nonuniform WENO divides ~30 times per face) run
// beside the bulk instead of after it.  In FP64 a bulk point whose stage
// field changes sign beside it takes the reference's nonuniform stencil: it
// is appended to `over`, counted in over_count[seq & 1], for launch 2; the
// other slot is cleared here for the next stage.  FP32 evaluates such a
// point in the bulk (defers_crossings) and has no launch 2.
template <int R, typename T, int S = IFACE_WENO>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(stage_waves<T>()))) k_reinit_stage(T* __restrict__ out, const T* __restrict__ fs,
                                                      const T* __restrict__ fn, const uint8_t* __restrict__ flags,
                                                      IList il, Box b, Inv invs, int stage, int* __restrict__ over,
                                                      int* __restrict__ over_count, int seq, int depth, int niface)
{
    __shared__ T lds[2][Shape<R>::N];
    const int blk = (int)blockIdx.x;
    if (blk == 0 && threadIdx.x == 0) over_count[(seq + 1) & 1] = 0;
    if (blk < niface) {
        const int total = *il.count;
        for (int j0 = 0; j0 < total; j0 += axis_lane_stride(niface)) {
            const AxisLane al = axis_lane(j0, blk);
            const bool active = al.point >= 0 && al.point < total;
            ReinitLane q;
            if (active) {
                const idx_t ind = (idx_t)il.idx[al.point];
                q = reinit_iface_lane<T, S>(fs, flags, b, invs, il, al.point, ind, ind, al.axis);
            }
            reinit_iface_finish(out, fs, fn, q, al, active, stage);
        }
        return;
    }
    int r0, c0, p0, p1;
    zmarch_share<R>(b, depth, blk - niface, (int)gridDim.x - niface, r0, c0, p0, p1);
    const idx_t plane = (idx_t)b.rows * b.cols;
    ZFace<T> zf;
    zmarch<R, T, false, true>(lds[0], lds[1], b, r0, c0, p0, p1, fs, fn, nullptr, nullptr, nullptr, flags, stage,
                        [&](const T* l, const ZCol<R>& zc, const T w[7], const ZStep<R, T, false>& st, int p) {
                            bool cross;
                            const T res = reinit_value<R>(l, zc.base, w, st.fn, st.flags(zc), invs, stage, cross,
                                                          zf, p == p0);
                            const bool bulk = !(st.flags(zc) & FLAG_SUBCELL);
                            if constexpr (defers_crossings<T>()) {
                                if (bulk && !cross) stp(out + (idx_t)p * plane, zc.in, res);
                                wave_append(bulk && cross, (int)((idx_t)p * plane + zc.in), over,
                                            &over_count[seq & 1]);
                            } else if (bulk) {
                                stp(out + (idx_t)p * plane, zc.in, res);
                            }
                        });
}

// launch 2 of 2: the deferred bulk points, three lanes per point, with the
// interface pass's arithmetic (spacing distances, the bulk pseudo-time step)
template <typename T, int S = IFACE_WENO>
__global__ void __launch_bounds__(256) k_reinit_deferred(T* __restrict__ out, const T* __restrict__ fs,
                                                         const T* __restrict__ fn, const uint8_t* __restrict__ flags,
                                                         IList il, Box b, Inv inv, int stage,
                                                         const int* __restrict__ over,
                                                         const int* __restrict__ over_count, int seq)
{
    const int total = over_count[seq & 1];
    for (int j0 = 0; j0 < total; j0 += axis_lane_stride()) {
        const AxisLane al = axis_lane(j0);
        const bool active = al.point >= 0 && al.point < total;
        ReinitLane q;
        if (active) {
            const idx_t ind = (idx_t)over[al.point];
            q = reinit_iface_lane<T, S>(fs, flags, b, inv, il, -1, ind, ind, al.axis);
        }
        reinit_iface_finish(out, fs, fn, q, al, active, stage);
    }
}

// ---------------------------------------------------------------------------
// interface lists: per-point data in list order, so the interface kernels
// read it coalesced
// ---------------------------------------------------------------------------
// boundary_interpolant (boundary_interpolant.cu:102-141) at the listed points
// only (it is zero elsewhere), one lane per (point, axis), in list order:
// cp[2a][j] = fore value, cp[2a+1][j] = back value
template <typename T>
__global__ void __launch_bounds__(256) k_list_interp(double* __restrict__ cp, IList il, const T* __restrict__ C,
                                                     Box b)
{
    const int total = *il.count;
    const double sp[3] = {b.dx, b.dy, b.dz};
    for (int j0 = 0; j0 < total; j0 += axis_lane_stride()) {
        const AxisLane al = axis_lane(j0);
        if (!(al.point >= 0 && al.point < total)) continue;
        const int a = al.axis, j = al.point;
        const double rf = il.dist[(size_t)(2 * a) * il.cap + j], lb = il.dist[(size_t)(2 * a + 1) * il.cap + j];
        Pair q{0.0, 0.0};
        if (rf < sp[a] || lb < sp[a]) {
            int r, c, p;
            coords_of(b, (idx_t)il.idx[j], r, c, p);
            double v[7];
            line7(C, r, c, p, a, b, v);
            q = sixth_interp(rf, lb, sp[a], v[0], v[1], v[2], v[3], v[4], v[5], v[6]);
        }
        cp[(size_t)(2 * a) * il.cap + j] = q.fore;
        cp[(size_t)(2 * a + 1) * il.cap + j] = q.back;
    }
}

// the cubic interface values of the WENO-RK3 extension (cubic_interp,
// wenork3_extend.cu:310-329) at the listed points, list order
__global__ void __launch_bounds__(256) k_list_interp_cubic(double* __restrict__ cp, IList il,
                                                           const double* __restrict__ C, Box b)
{
    const int total = *il.count;
    const double sp[3] = {b.dx, b.dy, b.dz};
    for (int j0 = 0; j0 < total; j0 += axis_lane_stride()) {
        const AxisLane al = axis_lane(j0);
        if (!(al.point >= 0 && al.point < total)) continue;
        const int a = al.axis, j = al.point;
        const double rf = il.dist[(size_t)(2 * a) * il.cap + j], lb = il.dist[(size_t)(2 * a + 1) * il.cap + j];
        double cf = 0.0, cb = 0.0;
        if (rf < sp[a] || lb < sp[a]) {
            int r, c, p;
            coords_of(b, (idx_t)il.idx[j], r, c, p);
            double v[7];
            line7(C, r, c, p, a, b, v);
            cubic_interface_values(cf, cb, rf, lb, sp[a], v);
        }
        cp[(size_t)(2 * a) * il.cap + j] = cf;
        cp[(size_t)(2 * a + 1) * il.cap + j] = cb;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_list_gather_vel(double* __restrict__ lv, IList il, T* __restrict__ vx,
                                                         const T* __restrict__ vy, const T* __restrict__ vz)
{
    const int cnt = *il.count;
    for (int j = (int)(blockIdx.x * 256 + threadIdx.x); j < cnt; j += (int)(gridDim.x * 256)) {
        const idx_t ind = (idx_t)il.idx[j];
        lv[j] = (double)vx[ind];
        lv[(size_t)il.cap + j] = (double)vy[ind];
        lv[(size_t)2 * il.cap + j] = (double)vz[ind];
        vx[ind] = T(__builtin_nan(""));  // the bulk of the extension stages skips the listed points
    }
}

// ---------------------------------------------------------------------------
// ENO2 reinitialization (1_0_ENORK2_Reinitialization/boundary_correction.cu:147-200)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double eno2_residual(const double* __restrict__ f, int r, int c, int p, const Box& b,
                                                const Dists& q, bool inside, double dt)
{
    const double sp[3] = {b.dx, b.dy, b.dz};
    double fo[3], bk[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7];
        line7(f, r, c, p, a, b, v);
        eno2_pair(fo[a], bk[a], v, q.r[a], q.l[a], sp[a]);
    }
    return godunov_step(inside, fo[0], bk[0], fo[1], bk[1], fo[2], bk[2], dt);
}

__global__ void __launch_bounds__(256) k_eno2_re_step(double* __restrict__ step, const double* __restrict__ lsf,
                                                      const uint8_t* __restrict__ mask,
                                                      const double* __restrict__ deltat, Dist6c d, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    Dists q;
#pragma unroll
    for (int a = 0; a < 3; ++a) { q.r[a] = d.r[a][ind]; q.l[a] = d.l[a][ind]; }
    step[ind] = eno2_residual(lsf, r, c, p, b, q, mask[ind] != 0, deltat[ind]);
}

// fused ENO2 stage.  rk: 0 = forward Euler (out = fs - L(fs)); 1 = Heun stage 1
// (out = fs - L(fs)); 2 = Heun stage 2 (out = (fn + fs - L(fs))/2,
// ENORK2Reinitialization.m:39).  dt = cfl*min_dist.
__global__ void __launch_bounds__(256) k_eno2_stage(double* __restrict__ out, const double* __restrict__ fs,
                                                    const double* __restrict__ fn, const uint8_t* __restrict__ flags,
                                                    Dist6c d, Box b, double cfl, int heun2)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const uint8_t fl = flags[ind];
    Dists q;
    if (fl & FLAG_SUBCELL) {
#pragma unroll
        for (int a = 0; a < 3; ++a) { q.r[a] = d.r[a][ind]; q.l[a] = d.l[a][ind]; }
    } else {
        q.r[0] = q.l[0] = b.dx; q.r[1] = q.l[1] = b.dy; q.r[2] = q.l[2] = b.dz;
    }
    const double step = eno2_residual(fs, r, c, p, b, q, (fl & FLAG_MASK) != 0, cfl * min6(q));
    out[ind] = heun2 ? (fn[ind] + fs[ind] - step) / 2 : fs[ind] - step;
}

// ---------------------------------------------------------------------------
// extension (2_2_WENO5RK3_Extend)
// ---------------------------------------------------------------------------

// upwind_normal: upwind_normal.cu:4-104
template <typename T>
__global__ void __launch_bounds__(256) k_upwind_normal(T* __restrict__ nx, T* __restrict__ ny,
                                                       T* __restrict__ nz, const T* __restrict__ lsf,
                                                       Dist6c d, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double sp[3] = {b.dx, b.dy, b.dz};
    T* out[3] = {nx, ny, nz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7], fo, bk;
        line7(lsf, r, c, p, a, b, v);
        extend_axis(fo, bk, v, d.r[a][ind], d.l[a][ind], sp[a], jpd(sp[a]), 0.0, 0.0);
        out[a][ind] = (T)((fabs(v[4]) < fabs(v[2])) ? fo : bk);
    }
}

// v = sign(F)*n/max(|n|,floor) in place on (vx,vy,vz) (WENO5RK3Extend.m:23-31)
template <typename T>
__global__ void k_extension_velocity(T* __restrict__ vx, T* __restrict__ vy, T* __restrict__ vz,
                                     const T* __restrict__ F, double floor_, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double fx = vx[i], fy = vy[i], fz = vz[i];
    const double g = dmax(sqrt(fx * fx + fy * fy + fz * fz), floor_);
    const double s = msign((double)F[i]);
    vx[i] = (T)(s * (fx / g));
    vy[i] = (T)(s * (fy / g));
    vz[i] = (T)(s * (fz / g));
}

// The extension velocity (WENO5RK3Extend.m:19-31) in two launches over
// disjoint points: per axis the component of upwind_normal.cu:4-104 (the
// fore derivative where |F(+1)| < |F(-1)|, else the back one; only that side
// is evaluated), rounded to T as upwind_normal stores it, then
// v = sign(F) n / max(|n|, floor) as k_extension_velocity computes it.  The
// first launch takes the points without a subcell distance (uniform WENO5 on
// every axis), the second the listed points (the nonuniform form on an axis
// with a subcell distance, read from the list).
__device__ __forceinline__ double upwind_normal_axis(const double v[7], double ds, bool fore)
{
    const JPd j = jpd(ds);
    const JPcd<double> m = jp_centre(v);
    return fore ? (m.c + jp_q(m.d[4], m.d[3], m.d[2], m.d[1], j.e)) * j.k
                : (m.c - jp_q(m.d[0], m.d[1], m.d[2], m.d[3], j.e)) * j.k;
}

template <typename T>
__device__ __forceinline__ void store_extension_velocity(T* __restrict__ vx, T* __restrict__ vy,
                                                         T* __restrict__ vz, idx_t ind, const double n[3], double f,
                                                         double floor_)
{
    const double g = dmax(sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]), floor_);
    const double sg = msign(f);
    vx[ind] = (T)(sg * (n[0] / g));
    vy[ind] = (T)(sg * (n[1] / g));
    vz[ind] = (T)(sg * (n[2] / g));
}

// the first launch, on the z-marching walk of the stage kernels: x- and
// y-taps from the staged tile, z-taps from the register window
template <int R, typename T>
__global__ void __launch_bounds__(256) k_extension_normal_z(T* __restrict__ vx, T* __restrict__ vy,
                                                            T* __restrict__ vz, const T* __restrict__ lsf,
                                                            const uint8_t* __restrict__ flags, double floor_, Box b,
                                                            int depth)
{
    __shared__ T lds[2][Shape<R>::N];
    int r0, c0, p0, p1;
    zmarch_share<R>(b, depth, (int)blockIdx.x, (int)gridDim.x, r0, c0, p0, p1);
    const idx_t plane = (idx_t)b.rows * b.cols;
    const double sp[3] = {b.dx, b.dy, b.dz};
    zmarch<R, T, false, true>(lds[0], lds[1], b, r0, c0, p0, p1, lsf, nullptr, nullptr, nullptr, nullptr, flags, 0,
                              [&](const T* l, const ZCol<R>& zc, const T w[7], const ZStep<R, T, false>& st, int p) {
                                  if (st.flags(zc) & FLAG_SUBCELL) return;  // listed, or outside the box
                                  double n[3];
#pragma unroll
                                  for (int a = 0; a < 3; ++a) {
                                      T t[7];
                                      if (a == 0) tile_line_x<R>(l, zc.base, t);
                                      else if (a == 1) tile_line_y(l, zc.base, t);
                                      else
#pragma unroll
                                          for (int k = 0; k < 7; ++k) t[k] = w[k];
                                      double v[7];
#pragma unroll
                                      for (int k = 0; k < 7; ++k) v[k] = (double)t[k];
                                      n[a] = (double)(T)upwind_normal_axis(v, sp[a], fabs(v[4]) < fabs(v[2]));
                                  }
                                  store_extension_velocity(vx, vy, vz, (idx_t)p * plane + zc.in, n, (double)w[3],
                                                           floor_);
                              });
}

template <typename T>
__global__ void __launch_bounds__(256) k_extension_normal_list(T* __restrict__ vx, T* __restrict__ vy,
                                                               T* __restrict__ vz, const T* __restrict__ lsf,
                                                               IList il, double floor_, Box b)
{
    const int cnt = *il.count;
    const double sp[3] = {b.dx, b.dy, b.dz};
    for (int j = (int)(blockIdx.x * 256 + threadIdx.x); j < cnt; j += (int)(gridDim.x * 256)) {
        const idx_t ind = (idx_t)il.idx[j];
        int r, c, p;
        coords_of(b, ind, r, c, p);
        double n[3];
#pragma unroll 1
        for (int a = 0; a < 3; ++a) {
            double v[7];
            line7(lsf, r, c, p, a, b, v);
            const bool fore = fabs(v[4]) < fabs(v[2]);
            const double ds = sp[a];
            const double r1 = il.dist[(size_t)(2 * a) * il.cap + j], l1 = il.dist[(size_t)(2 * a + 1) * il.cap + j];
            const double dv = (r1 < ds || l1 < ds)
                                  ? weno_nonuniform_one(v[0], v[1], v[2], v[3], v[4], v[5], v[6], r1, l1, ds, 0.0,
                                                        0.0, !fore)
                                  : upwind_normal_axis(v, ds, fore);
            n[a] = (double)(T)dv;
        }
        store_extension_velocity(vx, vy, vz, ind, n, (double)lsf[ind], floor_);
    }
}

// boundary_interpolant: boundary_interpolant.cu:102-141
__global__ void __launch_bounds__(256) k_boundary_interpolant(Dist6 cp, Dist6c d, const double* __restrict__ C,
                                                              Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double sp[3] = {b.dx, b.dy, b.dz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double rf = d.r[a][ind], lb = d.l[a][ind];
        Pair q{0.0, 0.0};
        if (rf < sp[a] || lb < sp[a]) {
            double v[7];
            line7(C, r, c, p, a, b, v);
            q = sixth_interp(rf, lb, sp[a], v[0], v[1], v[2], v[3], v[4], v[5], v[6]);
        }
        cp.r[a][ind] = q.fore;
        cp.l[a][ind] = q.back;
    }
}

__device__ __forceinline__ double extend_rate(const double* __restrict__ C, int r, int c, int p, idx_t ind,
                                              const Box& b, const Dists& q, const double vel[3], const Dist6c& cp,
                                              bool subcell)
{
    const double sp[3] = {b.dx, b.dy, b.dz};
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7], fo, bk;
        line7(C, r, c, p, a, b, v);
        double hf = 0.0, hb = 0.0;
        if (subcell) { hf = cp.r[a][ind]; hb = cp.l[a][ind]; }
        extend_axis(fo, bk, v, q.r[a], q.l[a], sp[a], jpd(sp[a]), hf, hb);
        // extend_step.cu:121-123 sums x, y, z terms left to right
        acc = acc + dmin(0.0, vel[a]) * fo;
        acc = acc + dmax(0.0, vel[a]) * bk;
    }
    return acc;
}

// drop-in extend_step: extend_step.cu:10-125
__global__ void __launch_bounds__(256) k_extend_step(double* __restrict__ step, const double* __restrict__ deltat,
                                                     const double* __restrict__ C, const double* __restrict__ vx,
                                                     const double* __restrict__ vy, const double* __restrict__ vz,
                                                     Dist6c d, Dist6c cp, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    Dists q;
#pragma unroll
    for (int a = 0; a < 3; ++a) { q.r[a] = d.r[a][ind]; q.l[a] = d.l[a][ind]; }
    if (frozen(q, b.dx)) { step[ind] = 0; return; }
    const double vel[3] = {vx[ind], vy[ind], vz[ind]};
    step[ind] = extend_rate(C, r, c, p, ind, b, q, vel, cp, true) * deltat[ind];
}

// Fused WENO5-RK3 extension stage, one launch: the first niface blocks walk
// the interface list with three lanes per point (see AxisLane; each lane
// evaluates the one upwind derivative its axis needs, on the nonuniform
// stencil where the axis crosses; lcp, lv: interface values and velocity in
// list order), the others z-march over every point without subcell
// distances.  The list blocks are dispatched first; their long dependent
// chains (the nonuniform WENO divides ~30 times) then run beside the bulk
// blocks instead of after them.  The two sets of points are disjoint.
template <int R, typename T, int S = IFACE_WENO>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(stage_waves<T>()))) k_extend_stage(T* __restrict__ out, const T* __restrict__ cs,
                                                      const T* __restrict__ cn, const T* __restrict__ vx,
                                                      const T* __restrict__ vy, const T* __restrict__ vz,
                                                      IList il, const double* __restrict__ lcp,
                                                      const double* __restrict__ lv, Box b, Inv inv, int stage,
                                                      int depth, int niface)
{
    __shared__ T lds[2][Shape<R>::N];
    const int blk = (int)blockIdx.x;
    if (blk < niface) {
        const int total = *il.count;
        for (int j0 = 0; j0 < total; j0 += axis_lane_stride(niface)) {
            const AxisLane al = axis_lane(j0, blk);
            const bool active = al.point >= 0 && al.point < total;
            ExtendLane q;
            if (active) {
                const idx_t ind = (idx_t)il.idx[al.point];
                q = extend_iface_lane<T, S>(cs, b, inv, il, lcp, lv, al.point, ind, ind, al.axis);
            }
            extend_iface_finish(out, cs, cn, q, al, active, stage);
        }
        return;
    }
    int r0, c0, p0, p1;
    zmarch_share<R>(b, depth, blk - niface, (int)gridDim.x - niface, r0, c0, p0, p1);
    const idx_t plane = (idx_t)b.rows * b.cols;
    // the listed points carry a NaN x-velocity (k_list_gather_vel), so the
    // bulk needs no flags: it skips them and the lanes outside the box
    zmarch<R, T, true, false>(lds[0], lds[1], b, r0, c0, p0, p1, cs, cn, vx, vy, vz, nullptr, stage,
                       [&](const T* l, const ZCol<R>& zc, const T w[7], const ZStep<R, T, true>& st, int p) {
                           const T res = upwind_value<R>(l, zc.base, w, st.vel, st.fn, (T)inv.dt_bulk, inv, stage);
                           if (zc.valid && st.vel[0] == st.vel[0]) stp(out + (idx_t)p * plane, zc.in, res);
                       });
}

// ---------------------------------------------------------------------------
// WENO5 advection operator (4_0_Calculus_ToolBox/advection_step.cu:30-56) and
// its six-output twin (5_0_Nonoscillating_Reconstruction/nonoscillating_interpolant.cu:59-122)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double advect_rate(const double* __restrict__ f, int r, int c, int p, const Box& b,
                                              const Inv& inv, const double vel[3])
{
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7];
        line7(f, r, c, p, a, b, v);
        acc = upwind_axis(acc, vel[a], v, inv.jd[a]);
    }
    return acc;
}

__global__ void __launch_bounds__(256) k_advection_step(double* __restrict__ astep, const double* __restrict__ vx,
                                                        const double* __restrict__ vy,
                                                        const double* __restrict__ vz,
                                                        const double* __restrict__ lsf, Box b, Inv inv)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double vel[3] = {vx[ind], vy[ind], vz[ind]};
    astep[ind] = advect_rate(lsf, r, c, p, b, inv, vel);
}

// fused TVD-RK3 advection stage: out = combine(fs - dt*L(fs), fn), plane tiles
template <int R, typename T>
__global__ void __launch_bounds__(256) k_advect_stage(T* __restrict__ out, const T* __restrict__ fs,
                                                      const T* __restrict__ fn, const T* __restrict__ vx,
                                                      const T* __restrict__ vy, const T* __restrict__ vz,
                                                      double dt, Box b, Inv inv, int stage, int depth)
{
    __shared__ T lds[2][Shape<R>::N];
    int r0, c0, p0, p1;
    zmarch_share<R>(b, depth, (int)blockIdx.x, (int)gridDim.x, r0, c0, p0, p1);
    const idx_t plane = (idx_t)b.rows * b.cols;
    zmarch<R, T, true, false>(lds[0], lds[1], b, r0, c0, p0, p1, fs, fn, vx, vy, vz, nullptr, stage,
                       [&](const T* l, const ZCol<R>& zc, const T w[7], const ZStep<R, T, true>& st, int p) {
                           const T res = upwind_value<R>(l, zc.base, w, st.vel, st.fn, (T)dt, inv, stage);
                           if (zc.valid) stp(out + (idx_t)p * plane, zc.in, res);
                       });
}

__global__ void __launch_bounds__(256) k_weno_derivative(Dist6 out, const double* __restrict__ lsf, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double sp[3] = {b.dx, b.dy, b.dz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7], fo, bk;
        line7(lsf, r, c, p, a, b, v);
        weno5_pair(fo, bk, v, jpd(sp[a]));
        out.l[a][ind] = bk;  // back_x, back_y, back_z
        out.r[a][ind] = fo;  // fore_x, fore_y, fore_z
    }
}

// ---------------------------------------------------------------------------
// calculus toolbox (4_0_Calculus_ToolBox/set_calculus_toolbox.cu:9-107)
// ---------------------------------------------------------------------------
struct Box27 {
    double v[3][3][3];  // [dz+1][dc+1][dr+1]
};

template <typename T>
__device__ __forceinline__ void load_box27(Box27& o, const T* __restrict__ f, int r, int c, int p, const Box& b)
{
#pragma unroll
    for (int k = -1; k <= 1; ++k) {
        const int pp = clampi(p + k, b.pges - 1);
#pragma unroll
        for (int j = -1; j <= 1; ++j) {
            const int cc = clampi(c + j, b.cols - 1);
#pragma unroll
            for (int i = -1; i <= 1; ++i) {
                const int rr = clampi(r + i, b.rows - 1);
                o.v[k + 1][j + 1][i + 1] = (double)f[(idx_t)pp * b.rows * b.cols + (idx_t)cc * b.rows + rr];
            }
        }
    }
}

struct Toolbox {
    double fx, fy, fz, g, fxx, fyy, fzz, fxy, fyz, fzx, lap;
};

// F_(dr,dc,dp) of the reference's neighbour names
#define T27(o, dr, dc, dp) (o).v[(dp) + 1][(dc) + 1][(dr) + 1]

__device__ __forceinline__ Toolbox toolbox_of(const Box27& F, const Box& b, double ds)
{
    Toolbox t;
    t.fx = (T27(F, 0, 1, 0) - T27(F, 0, -1, 0)) / (2 * b.dx);
    t.fy = (T27(F, 1, 0, 0) - T27(F, -1, 0, 0)) / (2 * b.dy);
    t.fz = (T27(F, 0, 0, 1) - T27(F, 0, 0, -1)) / (2 * b.dz);
    t.g = dmax(sqrt(t.fx * t.fx + t.fy * t.fy + t.fz * t.fz), 1e-14);
    const double f0 = T27(F, 0, 0, 0);
    t.fxx = (T27(F, 0, 1, 0) - 2 * f0 + T27(F, 0, -1, 0)) / (b.dx * b.dx);
    t.fyy = (T27(F, 1, 0, 0) - 2 * f0 + T27(F, -1, 0, 0)) / (b.dy * b.dy);
    t.fzz = (T27(F, 0, 0, 1) - 2 * f0 + T27(F, 0, 0, -1)) / (b.dz * b.dz);
    const double den = 4 * ds * ds;
    t.fxy = (T27(F, 1, 1, 0) + T27(F, -1, -1, 0) - T27(F, 1, -1, 0) - T27(F, -1, 1, 0)) / den;
    t.fyz = (T27(F, 1, 0, 1) + T27(F, -1, 0, -1) - T27(F, 1, 0, -1) - T27(F, -1, 0, 1)) / den;
    t.fzx = (T27(F, 0, 1, 1) + T27(F, 0, -1, -1) - T27(F, 0, 1, -1) - T27(F, 0, -1, 1)) / den;
    t.lap = t.fxx + t.fyy + t.fzz;
    return t;
}

// Towers primal-function Heaviside and delta (set_calculus_toolbox.cu:91-106)
__device__ __forceinline__ void towers_h_delta(double& H, double& delta, const Box27& P, const Toolbox& t,
                                               const Box& b)
{
    const double px = (T27(P, 0, 1, 0) - T27(P, 0, -1, 0)) / (2 * b.dx);
    const double py = (T27(P, 1, 0, 0) - T27(P, -1, 0, 0)) / (2 * b.dy);
    const double pz = (T27(P, 0, 0, 1) - T27(P, 0, 0, -1)) / (2 * b.dz);
    const double dot = px * t.fx + py * t.fy + pz * t.fz;
    const double g2 = t.g * t.g;
    H = dot / g2;
    const double p0 = T27(P, 0, 0, 0);
    const double pxx = (T27(P, 0, 1, 0) - 2 * p0 + T27(P, 0, -1, 0)) / (b.dx * b.dx);
    const double pyy = (T27(P, 1, 0, 0) - 2 * p0 + T27(P, -1, 0, 0)) / (b.dy * b.dy);
    const double pzz = (T27(P, 0, 0, 1) - 2 * p0 + T27(P, 0, 0, -1)) / (b.dz * b.dz);
    delta = (pxx + pyy + pzz) / g2 - dot * t.lap / (g2 * t.g * t.g);
}

__global__ void __launch_bounds__(256) k_set_calculus_toolbox(Fields18 o, const double* __restrict__ lsf,
                                                              const double* __restrict__ hp, Box b, double ds)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    Box27 F, P;
    load_box27(F, lsf, r, c, p, b);
    load_box27(P, hp, r, c, p, b);
    const Toolbox t = toolbox_of(F, b, ds);
    o.f[0][ind] = t.fx; o.f[1][ind] = t.fy; o.f[2][ind] = t.fz; o.f[3][ind] = t.g;
    o.f[4][ind] = t.fx / t.g; o.f[5][ind] = t.fy / t.g; o.f[6][ind] = t.fz / t.g;
    o.f[7][ind] = t.fxx; o.f[8][ind] = t.fyy; o.f[9][ind] = t.fzz;
    o.f[10][ind] = t.fxy; o.f[11][ind] = t.fyz; o.f[12][ind] = t.fzx; o.f[13][ind] = t.lap;
    double c1 = t.fxx * t.fx + t.fxy * t.fy + t.fzx * t.fz;
    double c2 = t.fxy * t.fx + t.fyy * t.fy + t.fyz * t.fz;
    double c3 = t.fzx * t.fx + t.fyz * t.fy + t.fzz * t.fz;
    const double g3 = t.g * t.g * t.g;
    o.f[14][ind] = -t.lap / t.g + (t.fx * c1 + t.fy * c2 + t.fz * c3) / g3;
    c1 = (t.fyy * t.fzz - t.fyz * t.fyz) * t.fx + (t.fzx * t.fyz - t.fxy * t.fzz) * t.fy
         + (t.fxy * t.fyz - t.fzx * t.fyy) * t.fz;
    c2 = (t.fyz * t.fzx - t.fxy * t.fzz) * t.fx + (t.fxx * t.fzz - t.fzx * t.fzx) * t.fy
         + (t.fzx * t.fxy - t.fxx * t.fyz) * t.fz;
    c3 = (t.fxy * t.fyz - t.fyy * t.fzx) * t.fx + (t.fzx * t.fxy - t.fxx * t.fyz) * t.fy
         + (t.fxx * t.fyy - t.fxy * t.fxy) * t.fz;
    o.f[15][ind] = (t.fx * c1 + t.fy * c2 + t.fz * c3) / (g3 * t.g);
    double H, delta;
    towers_h_delta(H, delta, P, t, b);
    o.f[16][ind] = H;
    o.f[17][ind] = delta;
}
#undef T27

// ---------------------------------------------------------------------------
// auxiliary-field geometry (4_0_Calculus_ToolBox/auxi_set_calculus_toolbox.cu:9-199,
// with norm and cross_product of shared_utilities.cup:7-19).  pow(g,k) is the
// product g*...*g, as in the toolbox above.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cross3(double& wx, double& wy, double& wz, double ux, double uy, double uz,
                                       double vx, double vy, double vz)
{
#pragma clang fp contract(off)
    wx = uy * vz - uz * vy;
    wy = uz * vx - ux * vz;
    wz = ux * vy - uy * vx;
}
__device__ __forceinline__ double norm3(double x, double y, double z)
{
    return dmax(sqrt(x * x + y * y + z * z), 1e-14);
}

__global__ void __launch_bounds__(256) k_auxi_set_calculus_toolbox(Fields27 o, const double* __restrict__ A,
                                                                   const double* __restrict__ ahp, Fields13c fin,
                                                                   Box b, double ds)
{
#pragma clang fp contract(off)
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    Box27 Ab, Pb;
    load_box27(Ab, A, r, c, p, b);
    load_box27(Pb, ahp, r, c, p, b);
    const Toolbox a = toolbox_of(Ab, b, ds);  // Ax..Azx, |grad A| = norm, ALaplacian
    o.f[0][ind] = a.fx; o.f[1][ind] = a.fy; o.f[2][ind] = a.fz; o.f[3][ind] = a.g;
    const double fx = fin.f[0][ind], fy = fin.f[1][ind], fz = fin.f[2][ind], fg = fin.f[3][ind];
    double cx, cy, cz;
    cross3(cx, cy, cz, fx, fy, fz, a.fx, a.fy, a.fz);
    const double ncr = norm3(cx, cy, cz);
    o.f[4][ind] = cx; o.f[5][ind] = cy; o.f[6][ind] = cz; o.f[7][ind] = ncr;
    const double tx = cx / ncr, ty = cy / ncr, tz = cz / ncr;
    o.f[8][ind] = tx; o.f[9][ind] = ty; o.f[10][ind] = tz;
    double nx, ny, nz;
    cross3(nx, ny, nz, tx, ty, tz, fin.f[4][ind], fin.f[5][ind], fin.f[6][ind]);
    o.f[11][ind] = nx; o.f[12][ind] = ny; o.f[13][ind] = nz;
    o.f[14][ind] = a.fxx; o.f[15][ind] = a.fyy; o.f[16][ind] = a.fzz;
    o.f[17][ind] = a.fxy; o.f[18][ind] = a.fyz; o.f[19][ind] = a.fzx; o.f[20][ind] = a.lap;
    const double fxx = fin.f[7][ind], fyy = fin.f[8][ind], fzz = fin.f[9][ind];
    const double fxy = fin.f[10][ind], fyz = fin.f[11][ind], fzx = fin.f[12][ind];
    // geodesic curvature (auxi_set_calculus_toolbox.cu:115-131)
    double vx = tx * fxx + ty * fxy + tz * fzx;
    double vy = tx * fxy + ty * fyy + tz * fyz;
    double vz = tx * fzx + ty * fyz + tz * fzz;
    double w1x, w1y, w1z;
    cross3(w1x, w1y, w1z, vx, vy, vz, a.fx, a.fy, a.fz);
    vx = tx * a.fxx + ty * a.fxy + tz * a.fzx;
    vy = tx * a.fxy + ty * a.fyy + tz * a.fyz;
    vz = tx * a.fzx + ty * a.fyz + tz * a.fzz;
    double w2x, w2y, w2z;
    cross3(w2x, w2y, w2z, fx, fy, fz, vx, vy, vz);
    o.f[21][ind] = (nx * (w1x + w2x) + ny * (w1y + w2y) + nz * (w1z + w2z)) / ncr;
    // the shape operator of F's level sets, dN/dx (:135-145)
    const double g3 = fg * fg * fg;
    const double sx = fxx * fx + fxy * fy + fzx * fz;
    const double sy = fxy * fx + fyy * fy + fyz * fz;
    const double sz = fzx * fx + fyz * fy + fzz * fz;
    const double Nxx = fxx / fg - fx * sx / g3, Nyx = fxy / fg - fy * sx / g3, Nzx = fzx / fg - fz * sx / g3;
    const double Nxy = fxy / fg - fx * sy / g3, Nyy = fyy / fg - fy * sy / g3, Nzy = fyz / fg - fz * sy / g3;
    const double Nxz = fzx / fg - fx * sz / g3, Nyz = fyz / fg - fy * sz / g3, Nzz = fzz / fg - fz * sz / g3;
    vx = Nxx * tx + Nxy * ty + Nxz * tz;
    vy = Nyx * tx + Nyy * ty + Nyz * tz;
    vz = Nzx * tx + Nzy * ty + Nzz * tz;
    o.f[22][ind] = -(tx * vx + ty * vy + tz * vz);
    vx = Nxx * nx + Nxy * ny + Nxz * nz;
    vy = Nyx * nx + Nyy * ny + Nyz * nz;
    vz = Nzx * nx + Nzy * ny + Nzz * nz;
    o.f[23][ind] = -(tx * vx + ty * vy + tz * vz);
    o.f[24][ind] = -(nx * vx + ny * vy + nz * vz);
    // Heaviside and delta of A from its primal (:166-197)
    double H, delta;
    towers_h_delta(H, delta, Pb, a, b);
    o.f[25][ind] = H;
    o.f[26][ind] = delta;
}

// periodic neighbour index of the @GD3 index maps (circshift, GD3.m:112-117)
__device__ __forceinline__ int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// GPUAsetCalculusToolBox.m:55-62: AFDiracDelta = C.(grad H x grad AH)/|C|^2 and
// ADiracDeltaDn = (Lap AH - ADiracDelta Lap A)/|grad A| with @GD3's periodic
// Fx/Fy/Fz (GD3.m:218-240) and Laplacian (:344-346, Fxx :248-250)
__global__ void __launch_bounds__(256) k_aux_toolbox_tail(double* __restrict__ afdd, double* __restrict__ addn,
                                                          const double* __restrict__ ah, const double* __restrict__ h,
                                                          const double* __restrict__ a, const double* __restrict__ cx,
                                                          const double* __restrict__ cy, const double* __restrict__ cz,
                                                          const double* __restrict__ ncross,
                                                          const double* __restrict__ add,
                                                          const double* __restrict__ agrad, Box b)
{
#pragma clang fp contract(off)
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const idx_t xp = lin(b, r, wrapi(c + 1, b.cols), p), xm = lin(b, r, wrapi(c - 1, b.cols), p);
    const idx_t yp = lin(b, wrapi(r + 1, b.rows), c, p), ym = lin(b, wrapi(r - 1, b.rows), c, p);
    const idx_t zp = lin(b, r, c, wrapi(p + 1, b.pges)), zm = lin(b, r, c, wrapi(p - 1, b.pges));
    const double h1x = (ah[xp] - ah[xm]) / (2 * b.dx), h1y = (ah[yp] - ah[ym]) / (2 * b.dy),
                 h1z = (ah[zp] - ah[zm]) / (2 * b.dz);
    const double h2x = (h[xp] - h[xm]) / (2 * b.dx), h2y = (h[yp] - h[ym]) / (2 * b.dy),
                 h2z = (h[zp] - h[zm]) / (2 * b.dz);
    double chx, chy, chz;
    cross3(chx, chy, chz, h2x, h2y, h2z, h1x, h1y, h1z);
    const double nc = ncross[ind];
    afdd[ind] = (cx[ind] * chx + cy[ind] * chy + cz[ind] * chz) / (nc * nc);
    const double lah = (ah[xp] - 2 * ah[ind] + ah[xm]) / (b.dx * b.dx) + (ah[yp] - 2 * ah[ind] + ah[ym]) / (b.dy * b.dy) +
                       (ah[zp] - 2 * ah[ind] + ah[zm]) / (b.dz * b.dz);
    const double la = (a[xp] - 2 * a[ind] + a[xm]) / (b.dx * b.dx) + (a[yp] - 2 * a[ind] + a[ym]) / (b.dy * b.dy) +
                      (a[zp] - 2 * a[ind] + a[zm]) / (b.dz * b.dz);
    addn[ind] = (lah - add[ind] * la) / agrad[ind];
}

// AHPrimal = max(A, 0) (GPUAsetCalculusToolBox.m:4); MATLAB's max ignores a NaN,
// so a NaN of A gives 0
__global__ void k_positive_part(double* __restrict__ out, const double* __restrict__ a, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] > 0.0 ? a[i] : 0.0;
}

// normal component of a velocity, u.N with N = grad F/max(|grad F|,1e-14) from
// clamped central differences (set_calculus_toolbox.cu:31-44)
template <typename T>
__global__ void __launch_bounds__(256) k_normal_velocity(T* __restrict__ out, const T* __restrict__ F,
                                                         const T* __restrict__ vx, const T* __restrict__ vy,
                                                         const T* __restrict__ vz, Box b)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double fx = ((double)F[lin(b, r, min(c + 1, b.cols - 1), p)] - (double)F[lin(b, r, max(c - 1, 0), p)]) /
                      (2 * b.dx);
    const double fy = ((double)F[lin(b, min(r + 1, b.rows - 1), c, p)] - (double)F[lin(b, max(r - 1, 0), c, p)]) /
                      (2 * b.dy);
    const double fz = ((double)F[lin(b, r, c, min(p + 1, b.pges - 1))] - (double)F[lin(b, r, c, max(p - 1, 0))]) /
                      (2 * b.dz);
    const double g = dmax(sqrt(fx * fx + fy * fy + fz * fz), 1e-14);
    out[ind] = (T)((double)vx[ind] * (fx / g) + (double)vy[ind] * (fy / g) + (double)vz[ind] * (fz / g));
}

// ---------------------------------------------------------------------------
// compensated reductions (area, volume, max|u|)
// ---------------------------------------------------------------------------

// cosine-smoothed delta/H of width eps (src/0_0_setup/+SD/@SDF3/SDF3.m:126-136);
// |grad F| by periodic central differences (src/0_0_setup/+SD/@GD3/GD3.m:212-223)
template <typename T>
__global__ void __launch_bounds__(256) k_area_volume_cosine(double* __restrict__ partial, const T* __restrict__ F,
                                                            Box b, double eps, idx_t n)
{
    KSum k[2] = {{0.0, 0.0}, {0.0, 0.0}};
    const double pi = 3.14159265358979323846;
    const idx_t stride = (idx_t)gridDim.x * blockDim.x;
    for (idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = (double)F[i];
        double delta = 0.0, heav;
        if (v > eps) {
            heav = 1.0;
        } else if (v < -eps) {
            heav = 0.0;
        } else {
            const idx_t plane = (idx_t)b.rows * b.cols;
            const int p = (int)(i / plane);
            const idx_t rem = i - (idx_t)p * plane;
            const int c = (int)(rem / b.rows);
            const int r = (int)(rem - (idx_t)c * b.rows);
            const int cp = (c + 1 == b.cols) ? 0 : c + 1, cm = (c == 0) ? b.cols - 1 : c - 1;
            const int rp = (r + 1 == b.rows) ? 0 : r + 1, rm = (r == 0) ? b.rows - 1 : r - 1;
            const int pp = (p + 1 == b.pges) ? 0 : p + 1, pm = (p == 0) ? b.pges - 1 : p - 1;
            const double fx = ((double)F[(idx_t)p * plane + (idx_t)cp * b.rows + r] -
                               (double)F[(idx_t)p * plane + (idx_t)cm * b.rows + r]) / (2 * b.dx);
            const double fy = ((double)F[(idx_t)p * plane + (idx_t)c * b.rows + rp] -
                               (double)F[(idx_t)p * plane + (idx_t)c * b.rows + rm]) / (2 * b.dy);
            const double fz = ((double)F[(idx_t)pp * plane + (idx_t)c * b.rows + r] -
                               (double)F[(idx_t)pm * plane + (idx_t)c * b.rows + r]) / (2 * b.dz);
            const double g = sqrt(fx * fx + fy * fy + fz * fz);
            delta = g * (1 + cos(pi * v / eps)) / eps / 2;
            heav = (1 + v / eps + sin(pi * v / eps) / pi) / 2;
        }
        k[0].add(delta);
        k[1].add(1 - heav);
    }
    block_reduce_store<2>(k, partial);
}

// Towers delta/H area and volume (+SD/@SDF3/SDF3.m:73-80) fused with the
// toolbox stencil: only points in the masks evaluate it.
template <typename T>
__global__ void __launch_bounds__(256) k_area_volume_towers(double* __restrict__ partial, const T* __restrict__ F,
                                                            Box b, double ds, idx_t n)
{
    KSum k[2] = {{0.0, 0.0}, {0.0, 0.0}};
    const idx_t stride = (idx_t)gridDim.x * blockDim.x;
    for (idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = (double)F[i];
        const bool in_area = fabs(v) < 2 * ds, in_vol = v < 2 * ds;
        if (!(in_area || in_vol)) continue;
        const idx_t plane = (idx_t)b.rows * b.cols;
        const int p = (int)(i / plane);
        const idx_t rem = i - (idx_t)p * plane;
        const int c = (int)(rem / b.rows);
        const int r = (int)(rem - (idx_t)c * b.rows);
        Box27 Fb, Pb;
        load_box27(Fb, F, r, c, p, b);
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int y = 0; y < 3; ++y)
#pragma unroll
                for (int z = 0; z < 3; ++z) Pb.v[x][y][z] = dmax(Fb.v[x][y][z], 0.0);
        const Toolbox t = toolbox_of(Fb, b, ds);
        double H, delta;
        towers_h_delta(H, delta, Pb, t, b);
        if (in_area) k[0].add(delta);
        if (in_vol) k[1].add(1 - H);
    }
    block_reduce_store<2>(k, partial);
}

// final: one block of 256 merges `nblk` partial (s,c) pairs per quantity in a
// fixed order; result[q] = s + c
template <int NQ>
__global__ void __launch_bounds__(256) k_reduce_final(double* __restrict__ result, const double* __restrict__ partial,
                                                      int nblk, double scale)
{
    KSum k[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) k[q] = {0.0, 0.0};
    for (int i = threadIdx.x; i < nblk; i += blockDim.x)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            KSum o{partial[(size_t)i * 2 * NQ + 2 * q], partial[(size_t)i * 2 * NQ + 2 * q + 1]};
            k[q].merge(o);
        }
    block_reduce<NQ>(k);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) result[q] = (k[q].s + k[q].c) * scale;
}

template <typename T>
__global__ void __launch_bounds__(256) k_max_abs3(double* __restrict__ partial, const T* __restrict__ a,
                                                  const T* __restrict__ b2, const T* __restrict__ c3, idx_t n)
{
    double m = 0.0;
    const idx_t stride = (idx_t)gridDim.x * blockDim.x;
    for (idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        m = dmax(m, dmax(fabs((double)a[i]), dmax(fabs((double)b2[i]), fabs((double)c3[i]))));
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) m = dmax(m, shfl_xor_d(m, s));
    __shared__ double lds[4];
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = dmax(dmax(lds[0], lds[1]), dmax(lds[2], lds[3]));
}

__global__ void k_max_final(double* __restrict__ result, const double* __restrict__ partial, int nblk)
{
    double m = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) m = dmax(m, partial[i]);
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) m = dmax(m, shfl_xor_d(m, s));
    __shared__ double lds[4];
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) result[0] = dmax(dmax(lds[0], lds[1]), dmax(lds[2], lds[3]));
}

template <typename T>
__global__ void k_copy(T* __restrict__ dst, const T* __restrict__ src, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// dst = src converted (FP32 -> FP64 exactly, FP64 -> FP32 rounded to nearest)
template <typename D, typename S>
__global__ void k_convert(D* __restrict__ dst, const S* __restrict__ src, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (D)src[i];
}

template __global__ void k_reduce_final<2>(double*, const double*, int, double);

}  // namespace vls

// ===========================================================================
// host launchers
// ===========================================================================
namespace vls {

static inline Box box_of(const Geom& g) { return Box{g.rows, g.cols, g.pges, g.dx, g.dy, g.dz}; }

static inline Inv inv_of(const Geom& g)
{
    const double m = std::min(std::min(g.dx, g.dx), std::min(std::min(g.dy, g.dy), std::min(g.dz, g.dz)));
    Inv v{{1.0 / g.dx, 1.0 / g.dy, 1.0 / g.dz}, 0.3 * m, {}, {}};
    set_jp(v.jd, v.jf, g.dx, g.dy, g.dz);
    return v;
}

// one thread per point: 64 rows x 4 cols per block, one z-page per grid.z
static inline void point_grid(const Geom& g, dim3& grid, dim3& block)
{
    block = dim3(64, 4, 1);
    grid = dim3((g.rows + 63) / 64, (g.cols + 3) / 4, g.pges);
}

static inline dim3 flat_grid(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

#define VLS_POINT_LAUNCH(kernel, ...)                                   \
    do {                                                                \
        dim3 grid_, block_;                                             \
        point_grid(g, grid_, block_);                                   \
        hipLaunchKernelGGL(kernel, grid_, block_, 0, s, __VA_ARGS__);   \
        return hipGetLastError();                                       \
    } while (0)

template <typename T>
hipError_t launch_boundary_location(int order, const Dist6& d, const T* lsf, const Geom& g, hipStream_t s)
{
    dim3 grid, block;
    point_grid(g, grid, block);
    if (order == 2) k_boundary_location<2, T><<<grid, block, 0, s>>>(d, lsf, box_of(g));
    else if (order == 3) k_boundary_location<3, T><<<grid, block, 0, s>>>(d, lsf, box_of(g));
    else k_boundary_location<6, T><<<grid, block, 0, s>>>(d, lsf, box_of(g));
    return hipGetLastError();
}

static inline int bulk_blocks(const Geom& g, int depth);

template <typename T>
hipError_t launch_set_distance(int order, const Dist6& d, double* min_dist, uint8_t* flags, const T* mask_src,
                               const T* lsf, const ListView& lv, int* cand, int* cand_count, const Geom& g,
                               hipStream_t s)
{
    dim3 grid, block;
    point_grid(g, grid, block);
    k_zero_pair<<<dim3(1), dim3(64), 0, s>>>(cand_count, lv.idx ? lv.count : nullptr);
    if (!mask_src || (const void*)mask_src == (const void*)lsf) {
        const int zd = zmarch_depth(g), m = mask_src ? 1 : 0;
        if (g.tr == 32)
            k_crossings_z<32, T><<<dim3(bulk_blocks(g, zd)), dim3(256), 0, s>>>(d, min_dist, flags, m, lsf, box_of(g),
                                                                               zd, cand, cand_count);
        else
            k_crossings_z<64, T><<<dim3(bulk_blocks(g, zd)), dim3(256), 0, s>>>(d, min_dist, flags, m, lsf, box_of(g),
                                                                               zd, cand, cand_count);
    } else {
        k_crossings<T><<<grid, block, 0, s>>>(d, min_dist, flags, mask_src, lsf, box_of(g), cand, cand_count);
    }
    const dim3 cg(iface_blocks(g));
    if (order == 2)
        k_crossing_distances<2, T><<<cg, dim3(256), 0, s>>>(d, min_dist, flags, lsf, box_of(g), cand, cand_count, lv);
    else if (order == 3)
        k_crossing_distances<3, T><<<cg, dim3(256), 0, s>>>(d, min_dist, flags, lsf, box_of(g), cand, cand_count, lv);
    else
        k_crossing_distances<6, T><<<cg, dim3(256), 0, s>>>(d, min_dist, flags, lsf, box_of(g), cand, cand_count, lv);
    return hipGetLastError();
}

hipError_t launch_fill_dense(const Dist6& d, double* min_dist, const ListView& lv, const Geom& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_dense, flat_grid(g.n()), dim3(256), 0, s, d, min_dist, box_of(g), (idx_t)g.n());
    const IList il{lv.idx, lv.count, lv.dist, lv.dt, lv.cap};
    hipLaunchKernelGGL(k_scatter_list, dim3(iface_blocks(g)), dim3(256), 0, s, d, min_dist, il);
    return hipGetLastError();
}

hipError_t launch_zero_ints(int* a, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, a, n);
    return hipGetLastError();
}

// blocks of an interface-list launch (84 points per block of 256, three lanes
// per point): one sweep covers lists up to 1/48 of the box (a closed surface
// in a box of n^3 points lists O(n^2) of them); longer lists are covered by
// striding
int iface_blocks(const Geom& g)
{
    const int64_t b = g.n() / (84 * 48);
    return (int)(b < 16 ? 16 : (b > 4096 ? 4096 : b));
}

// compute units of the current device, read once
static int device_cus()
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipDeviceProp_t prop;
        cus = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
               prop.multiProcessorCount > 0)
                  ? prop.multiProcessorCount
                  : 256;
    }
    return cus;
}

// interface-list blocks of a fused stage launch: one per CU at most, so the
// list lanes (long dependent chains) hold a fifth to a quarter of the block
// slots and the bulk blocks the rest
static int fused_iface_blocks(const Geom& g)
{
    const int b = iface_blocks(g), c = device_cus();
    return b < c ? b : c;
}

// blocks of the deferred-point launch: the deferred points are few (points
// the interface reached during the reinitialization), a tenth of a list
int deferred_blocks(const Geom& g)
{
    const int b = iface_blocks(g) / 10;
    return b < 8 ? 8 : b;
}

// pages per block of the z-marching bulk stage launches: g.zd, or by
// default the largest power of two up to 16 (and pges) that still gives at
// least 2048 blocks (8 per CU of a 256-CU device); measured on MI355X,
// depths 4-16 are within 3% of the best at 128^3 and 256^3
int zmarch_depth(const Geom& g)
{
    if (g.zd > 0) return g.zd < g.pges ? g.zd : g.pges;
    const int64_t cols_ = g.tr == 32 ? (int64_t)tiles_per_col<32>(g.rows) * tiles_per_row<32>(g.cols)
                                     : (int64_t)tiles_per_col<64>(g.rows) * tiles_per_row<64>(g.cols);
    int d = 1;
    while (2 * d <= 16 && 2 * d <= g.pges && cols_ * g.pges / (2 * d) >= 2048) d *= 2;
    return d;
}

static inline int bulk_blocks(const Geom& g, int depth)
{
    return g.tr == 32 ? zmarch_blocks<32>(g.rows, g.cols, g.pges, depth)
                      : zmarch_blocks<64>(g.rows, g.cols, g.pges, depth);
}

template <typename T>
hipError_t launch_set_mask_bit(uint8_t* flags, const T* F, const Geom& g, hipStream_t s)
{
    k_set_mask_bit<T><<<flat_grid(g.n()), dim3(256), 0, s>>>(flags, F, (idx_t)g.n());
    return hipGetLastError();
}

hipError_t launch_re_step(double* step, const double* lsf, const uint8_t* mask, const double* deltat,
                          const Dist6c& d, const Geom& g, hipStream_t s)
{
    VLS_POINT_LAUNCH(k_re_step, step, lsf, mask, deltat, d, box_of(g));
}

hipError_t launch_eno2_re_step(double* step, const double* lsf, const uint8_t* mask, const double* deltat,
                               const Dist6c& d, const Geom& g, hipStream_t s)
{
    VLS_POINT_LAUNCH(k_eno2_re_step, step, lsf, mask, deltat, d, box_of(g));
}

// bulk, then the interface list and the deferred bulk points (FP64), in stream order
template <typename T, int S>
static hipError_t reinit_stage_scheme(T* out, const T* fs, const T* fn, const uint8_t* flags, const ListView& lv,
                                      const Geom& g, int stage, int* over, int* over_count, int seq, hipStream_t s)
{
    const IList il{lv.idx, lv.count, lv.dist, lv.dt, lv.cap};
    const int zd = zmarch_depth(g), ni = fused_iface_blocks(g);
    const dim3 grid(bulk_blocks(g, zd) + ni);
    if (g.tr == 32)
        k_reinit_stage<32, T, S><<<grid, dim3(256), 0, s>>>(out, fs, fn, flags, il, box_of(g), inv_of(g), stage,
                                                            over, over_count, seq, zd, ni);
    else
        k_reinit_stage<64, T, S><<<grid, dim3(256), 0, s>>>(out, fs, fn, flags, il, box_of(g), inv_of(g), stage,
                                                            over, over_count, seq, zd, ni);
    if constexpr (defers_crossings<T>())
        k_reinit_deferred<T, S><<<dim3(deferred_blocks(g)), dim3(256), 0, s>>>(out, fs, fn, flags, il, box_of(g),
                                                                                inv_of(g), stage, over, over_count,
                                                                                seq);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_reinit_stage(T* out, const T* fs, const T* fn, const uint8_t* flags, const ListView& lv,
                               const Geom& g, int stage, int* over, int* over_count, int seq, hipStream_t s)
{
    return reinit_stage_scheme<T, IFACE_WENO>(out, fs, fn, flags, lv, g, stage, over, over_count, seq, s);
}

hipError_t launch_cubic_reinit_stage(double* out, const double* fs, const double* fn, const uint8_t* flags,
                                     const ListView& lv, const Geom& g, int stage, int* over, int* over_count,
                                     int seq, hipStream_t s)
{
    return reinit_stage_scheme<double, IFACE_CUBIC>(out, fs, fn, flags, lv, g, stage, over, over_count, seq, s);
}

template <typename T>
hipError_t launch_list_interp(double* lcp, const ListView& lv, const T* C, const Geom& g, hipStream_t s)
{
    const IList il{lv.idx, lv.count, lv.dist, lv.dt, lv.cap};
    k_list_interp<T><<<dim3(iface_blocks(g)), dim3(256), 0, s>>>(lcp, il, C, box_of(g));
    return hipGetLastError();
}

hipError_t launch_list_interp_cubic(double* lcp, const ListView& lv, const double* C, const Geom& g, hipStream_t s)
{
    const IList il{lv.idx, lv.count, lv.dist, lv.dt, lv.cap};
    k_list_interp_cubic<<<dim3(iface_blocks(g)), dim3(256), 0, s>>>(lcp, il, C, box_of(g));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_list_gather_vel(double* lvel, const ListView& lv, T* vx, const T* vy, const T* vz,
                                  const Geom& g, hipStream_t s)
{
    const IList il{lv.idx, lv.count, lv.dist, lv.dt, lv.cap};
    k_list_gather_vel<T><<<dim3(iface_blocks(g)), dim3(256), 0, s>>>(lvel, il, vx, vy, vz);
    return hipGetLastError();
}

hipError_t launch_eno2_stage(double* out, const double* fs, const double* fn, const uint8_t* flags,
                             const Dist6c& d, const Geom& g, double cfl, int heun2, hipStream_t s)
{
    VLS_POINT_LAUNCH(k_eno2_stage, out, fs, fn, flags, d, box_of(g), cfl, heun2);
}

template <typename T>
hipError_t launch_upwind_normal(T* nx, T* ny, T* nz, const T* lsf, const Dist6c& d, const Geom& g, hipStream_t s)
{
    dim3 grid, block;
    point_grid(g, grid, block);
    k_upwind_normal<T><<<grid, block, 0, s>>>(nx, ny, nz, lsf, d, box_of(g));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_extension_normal(T* vx, T* vy, T* vz, const T* F, const uint8_t* flags, const ListView& lv,
                                   double floor_, const Geom& g, hipStream_t s)
{
    const int zd = zmarch_depth(g);
    if (g.tr == 32)
        k_extension_normal_z<32, T><<<dim3(bulk_blocks(g, zd)), dim3(256), 0, s>>>(vx, vy, vz, F, flags, floor_,
                                                                                  box_of(g), zd);
    else
        k_extension_normal_z<64, T><<<dim3(bulk_blocks(g, zd)), dim3(256), 0, s>>>(vx, vy, vz, F, flags, floor_,
                                                                                  box_of(g), zd);
    const IList il{lv.idx, lv.count, lv.dist, lv.dt, lv.cap};
    k_extension_normal_list<T><<<dim3(iface_blocks(g)), dim3(256), 0, s>>>(vx, vy, vz, F, il, floor_, box_of(g));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_extension_velocity(T* vx, T* vy, T* vz, const T* F, double floor_, const Geom& g, hipStream_t s)
{
    k_extension_velocity<T><<<flat_grid(g.n()), dim3(256), 0, s>>>(vx, vy, vz, F, floor_, (idx_t)g.n());
    return hipGetLastError();
}

hipError_t launch_boundary_interpolant(const Dist6& cp, const Dist6c& d, const double* C, const Geom& g,
                                       hipStream_t s)
{
    VLS_POINT_LAUNCH(k_boundary_interpolant, cp, d, C, box_of(g));
}

hipError_t launch_extend_step(double* step, const double* deltat, const double* C, const double* vx,
                              const double* vy, const double* vz, const Dist6c& d, const Dist6c& cp,
                              const Geom& g, hipStream_t s)
{
    VLS_POINT_LAUNCH(k_extend_step, step, deltat, C, vx, vy, vz, d, cp, box_of(g));
}

template <typename T, int S>
static hipError_t extend_stage_scheme(T* out, const T* cs, const T* cn, const T* vx, const T* vy, const T* vz,
                                      const ListView& lv, const double* lcp, const double* lvel, const Geom& g,
                                      int stage, hipStream_t s)
{
    const IList il{lv.idx, lv.count, lv.dist, lv.dt, lv.cap};
    const int zd = zmarch_depth(g), ni = fused_iface_blocks(g);
    const dim3 grid(bulk_blocks(g, zd) + ni);
    if (g.tr == 32)
        k_extend_stage<32, T, S><<<grid, dim3(256), 0, s>>>(out, cs, cn, vx, vy, vz, il, lcp, lvel,
                                                            box_of(g), inv_of(g), stage, zd, ni);
    else
        k_extend_stage<64, T, S><<<grid, dim3(256), 0, s>>>(out, cs, cn, vx, vy, vz, il, lcp, lvel,
                                                            box_of(g), inv_of(g), stage, zd, ni);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_extend_stage(T* out, const T* cs, const T* cn, const T* vx, const T* vy, const T* vz,
                               const ListView& lv, const double* lcp, const double* lvel, const Geom& g, int stage,
                               hipStream_t s)
{
    return extend_stage_scheme<T, IFACE_WENO>(out, cs, cn, vx, vy, vz, lv, lcp, lvel, g, stage, s);
}

hipError_t launch_cubic_extend_stage(double* out, const double* cs, const double* cn, const double* vx,
                                     const double* vy, const double* vz, const ListView& lv, const double* lcp,
                                     const double* lvel, const Geom& g, int stage, hipStream_t s)
{
    return extend_stage_scheme<double, IFACE_CUBIC>(out, cs, cn, vx, vy, vz, lv, lcp, lvel, g, stage, s);
}

hipError_t launch_advection_step(double* astep, const double* vx, const double* vy, const double* vz,
                                 const double* lsf, const Geom& g, hipStream_t s)
{
    VLS_POINT_LAUNCH(k_advection_step, astep, vx, vy, vz, lsf, box_of(g), inv_of(g));
}

template <typename T>
hipError_t launch_advect_stage(T* out, const T* fs, const T* fn, const T* vx, const T* vy, const T* vz, double dt,
                               const Geom& g, int stage, hipStream_t s)
{
    const int zd = zmarch_depth(g);
    if (g.tr == 32)
        k_advect_stage<32, T><<<dim3(bulk_blocks(g, zd)), dim3(256), 0, s>>>(out, fs, fn, vx, vy, vz, dt, box_of(g),
                                                                              inv_of(g), stage, zd);
    else
        k_advect_stage<64, T><<<dim3(bulk_blocks(g, zd)), dim3(256), 0, s>>>(out, fs, fn, vx, vy, vz, dt, box_of(g),
                                                                              inv_of(g), stage, zd);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_normal_velocity(T* out, const T* F, const T* vx, const T* vy, const T* vz, const Geom& g,
                                  hipStream_t s)
{
    dim3 grid, block;
    point_grid(g, grid, block);
    k_normal_velocity<T><<<grid, block, 0, s>>>(out, F, vx, vy, vz, box_of(g));
    return hipGetLastError();
}

hipError_t launch_weno_derivative(const Dist6& out, const double* lsf, const Geom& g, hipStream_t s)
{
    VLS_POINT_LAUNCH(k_weno_derivative, out, lsf, box_of(g));
}

hipError_t launch_set_calculus_toolbox(const Fields18& o, const double* lsf, const double* hp, const Geom& g,
                                       double ds, hipStream_t s)
{
    VLS_POINT_LAUNCH(k_set_calculus_toolbox, o, lsf, hp, box_of(g), ds);
}

hipError_t launch_auxi_set_calculus_toolbox(const Fields27& o, const double* lsf, const double* ahp,
                                            const Fields13c& fin, const Geom& g, double ds, hipStream_t s)
{
    VLS_POINT_LAUNCH(k_auxi_set_calculus_toolbox, o, lsf, ahp, fin, box_of(g), ds);
}

hipError_t launch_aux_toolbox_tail(double* afdd, double* addn, const double* ah, const double* h, const double* a,
                                   const double* cx, const double* cy, const double* cz, const double* ncross,
                                   const double* add, const double* agrad, const Geom& g, hipStream_t s)
{
    VLS_POINT_LAUNCH(k_aux_toolbox_tail, afdd, addn, ah, h, a, cx, cy, cz, ncross, add, agrad, box_of(g));
}

hipError_t launch_positive_part(double* out, const double* a, int64_t n, hipStream_t s)
{
    k_positive_part<<<flat_grid(n), dim3(256), 0, s>>>(out, a, (idx_t)n);
    return hipGetLastError();
}

// 1024 blocks x 256 threads: 4 blocks per CU on 256 CUs
int reduce_blocks() { return 1024; }

static inline int blocks_for(int64_t n)
{
    const int64_t want = (n + 255) / 256;
    return (int)(want < reduce_blocks() ? (want < 1 ? 1 : want) : reduce_blocks());
}

template <typename T>
hipError_t launch_area_volume_cosine(double* result, double* partial, const T* F, const Geom& g, double eps,
                                     double scale, hipStream_t s)
{
    const int nb = blocks_for(g.n());
    k_area_volume_cosine<T><<<dim3(nb), dim3(256), 0, s>>>(partial, F, box_of(g), eps, (idx_t)g.n());
    hipLaunchKernelGGL(k_reduce_final<2>, dim3(1), dim3(256), 0, s, result, partial, nb, scale);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_area_volume_towers(double* result, double* partial, const T* F, const Geom& g, double ds,
                                     double scale, hipStream_t s)
{
    const int nb = blocks_for(g.n());
    k_area_volume_towers<T><<<dim3(nb), dim3(256), 0, s>>>(partial, F, box_of(g), ds, (idx_t)g.n());
    hipLaunchKernelGGL(k_reduce_final<2>, dim3(1), dim3(256), 0, s, result, partial, nb, scale);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_max_abs3(double* result, double* partial, const T* a, const T* b, const T* c, const Geom& g,
                           hipStream_t s)
{
    const int nb = blocks_for(g.n());
    k_max_abs3<T><<<dim3(nb), dim3(256), 0, s>>>(partial, a, b, c, (idx_t)g.n());
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, s, result, partial, nb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_copy(T* dst, const T* src, int64_t n, hipStream_t s)
{
    k_copy<T><<<flat_grid(n), dim3(256), 0, s>>>(dst, src, (idx_t)n);
    return hipGetLastError();
}

hipError_t launch_widen(double* dst, const float* src, int64_t n, hipStream_t s)
{
    k_convert<double, float><<<flat_grid(n), dim3(256), 0, s>>>(dst, src, (idx_t)n);
    return hipGetLastError();
}

hipError_t launch_narrow(float* dst, const double* src, int64_t n, hipStream_t s)
{
    k_convert<float, double><<<flat_grid(n), dim3(256), 0, s>>>(dst, src, (idx_t)n);
    return hipGetLastError();
}

// instantiations for FP64 fields and FP32 (bulk) fields
#define VLS_INSTANTIATE(T)                                                                                          \
    template hipError_t launch_boundary_location<T>(int, const Dist6&, const T*, const Geom&, hipStream_t);         \
    template hipError_t launch_set_mask_bit<T>(uint8_t*, const T*, const Geom&, hipStream_t);                       \
    template hipError_t launch_set_distance<T>(int, const Dist6&, double*, uint8_t*, const T*, const T*,            \
                                               const ListView&, int*, int*, const Geom&, hipStream_t);              \
    template hipError_t launch_reinit_stage<T>(T*, const T*, const T*, const uint8_t*, const ListView&, const Geom&, \
                                               int, int*, int*, int, hipStream_t);                                  \
    template hipError_t launch_list_interp<T>(double*, const ListView&, const T*, const Geom&, hipStream_t);        \
    template hipError_t launch_list_gather_vel<T>(double*, const ListView&, T*, const T*, const T*,                \
                                                  const Geom&, hipStream_t);                                        \
    template hipError_t launch_upwind_normal<T>(T*, T*, T*, const T*, const Dist6c&, const Geom&, hipStream_t);     \
    template hipError_t launch_extension_velocity<T>(T*, T*, T*, const T*, double, const Geom&, hipStream_t);       \
    template hipError_t launch_extension_normal<T>(T*, T*, T*, const T*, const uint8_t*, const ListView&, double,   \
                                                   const Geom&, hipStream_t);                                       \
    template hipError_t launch_extend_stage<T>(T*, const T*, const T*, const T*, const T*, const T*,               \
                                               const ListView&, const double*, const double*, const Geom&, int,      \
                                               hipStream_t);                                                        \
    template hipError_t launch_advect_stage<T>(T*, const T*, const T*, const T*, const T*, const T*, double,        \
                                               const Geom&, int, hipStream_t);                                      \
    template hipError_t launch_normal_velocity<T>(T*, const T*, const T*, const T*, const T*, const Geom&,          \
                                                  hipStream_t);                                                     \
    template hipError_t launch_area_volume_cosine<T>(double*, double*, const T*, const Geom&, double, double,       \
                                                     hipStream_t);                                                  \
    template hipError_t launch_area_volume_towers<T>(double*, double*, const T*, const Geom&, double, double,       \
                                                     hipStream_t);                                                  \
    template hipError_t launch_max_abs3<T>(double*, double*, const T*, const T*, const T*, const Geom&, hipStream_t); \
    template hipError_t launch_copy<T>(T*, const T*, int64_t, hipStream_t);

VLS_INSTANTIATE(double)
VLS_INSTANTIATE(float)
#undef VLS_INSTANTIATE

}  // namespace vls
