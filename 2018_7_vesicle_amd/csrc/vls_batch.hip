// Batched kernels: many grids of different sizes advanced by one launch per
// operator stage.
//
// A batch packs every field of all its grids back to back (each grid in the
// column-major layout of the single-grid kernels) and lists the plane tiles
// of all grids (64 x 4 or 32 x 8 points of one page, one shape per batch) in
// one array; tile_grid[] names each tile's grid.  The
// bulk kernels are persistent: a fixed number of blocks (a multiple of 8)
// walks the tile list, each of the 8 XCDs taking one contiguous eighth of it
// so that neighbouring tiles share their halos in one L2.  Interface lists
// hold packed element indices; a lane finds its grid by binary search over
// the grid offsets.  Every per-point body is the single-grid kernels' own
// (vls_tile.hpp), called with the pointers of the point's grid, so each grid
// of a batch gets the single-grid result bit for bit.
#include <type_traits>

#include "vls_tile.hpp"

namespace vls {

__device__ __forceinline__ Box box_of(const BGrid& G) { return Box{G.rows, G.cols, G.pges, G.dx, G.dy, G.dz}; }

__device__ __forceinline__ Inv inv_of(const BGrid& G)
{
    Inv v;
    v.i[0] = G.inv[0];
    v.i[1] = G.inv[1];
    v.i[2] = G.inv[2];
    v.dt_bulk = G.dt_bulk;
    for (int a = 0; a < 3; ++a) {
        v.jd[a] = G.jd[a];
        v.jf[a] = G.jf[a];
    }
    return v;
}

// grid of a packed element index: offs[g] <= gind < offs[g+1]
__device__ __forceinline__ int grid_of(const BatchView& B, int64_t gind)
{
    int lo = 0, hi = B.ngrids;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (B.offs[mid] <= gind) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <typename T>
__device__ __forceinline__ const T* shift(const T* p, int64_t off)
{
    return p ? p + off : p;
}
template <typename T>
__device__ __forceinline__ T* shift(T* p, int64_t off)
{
    return p ? p + off : p;
}

// ---------------------------------------------------------------------------
// list building: setDistance (setDistance.m:5-31 with boundary_location.cu:84-132)
// and the interface list, in two passes as for one grid (k_crossings,
// k_crossing_distances): every point tests its six neighbours for a crossing
// and writes its flags (mask bit only), the points with a crossing go to a
// candidate list (block-aggregated); one lane per candidate then computes its
// six distances (the r-side from the crossing (i, i+1) on the stencil around
// i, the l-side as spacing minus the distance of the crossing (i-1, i) on the
// stencil around i-1, which is what boundary_location writes into the
// l-side array of i) and, where one differs from the spacing, sets
// FLAG_SUBCELL and appends the point to the list with its distances, time
// step and grid.  No full distance arrays are stored.
// ---------------------------------------------------------------------------
template <int R, typename T>
__global__ void __launch_bounds__(256) kb_crossings(const T* __restrict__ F, const T* __restrict__ mask_src,
                                                    uint8_t* __restrict__ flags, BatchView B, int* __restrict__ cand,
                                                    int* __restrict__ cand_count)
{
    __shared__ int cbuf[1024], ccount, cbase;
    BlockList<1024> bl{cbuf, &ccount, &cbase};
    bl.init();
    const TileRange tr = tile_range(B.ntiles);
    for (int tile = tr.first; tile < tr.end; tile += tr.step) {
        const int g = B.tile_grid[tile];
        const BGrid G = B.grids[g];
        const Box b = box_of(G);
        const Tile t = tile_at<R>(b, tile - G.tile0);
        const T* f = F + G.off;
        bool cross = false;
        idx_t ind = 0;
        if (t.valid) {
            ind = lin(b, t.r, t.c, t.p);
            const double f0 = (double)f[ind];
            const double xu = (double)f[lin(b, t.r, min(t.c + 1, b.cols - 1), t.p)];
            const double xd = (double)f[lin(b, t.r, max(t.c - 1, 0), t.p)];
            const double yu = (double)f[lin(b, min(t.r + 1, b.rows - 1), t.c, t.p)];
            const double yd = (double)f[lin(b, max(t.r - 1, 0), t.c, t.p)];
            const double zu = (double)f[lin(b, t.r, t.c, min(t.p + 1, b.pges - 1))];
            const double zd = (double)f[lin(b, t.r, t.c, max(t.p - 1, 0))];
            cross = (f0 * xu < 0) | (xd * f0 < 0) | (f0 * yu < 0) | (yd * f0 < 0) | (f0 * zu < 0) | (zd * f0 < 0);
            flags[G.off + ind] = (mask_src && mask_src[G.off + ind] < 0) ? (uint8_t)FLAG_MASK : (uint8_t)0;
        }
        bl.push(cross, (int)(G.off + ind), cand, cand_count);
    }
    bl.flush(cand, cand_count);
}

// ORDER: the locator, 6 (boundary_location), 2 (the quadratic-ENO
// boundary_correction of the ENO-RK2 variants) or 3 (the cubic one of the
// WENO-RK3 variants).
template <typename T, int ORDER>
__global__ void __launch_bounds__(256) kb_crossing_distances(const T* __restrict__ F, uint8_t* __restrict__ flags,
                                                             ListView il, BatchView B, const int* __restrict__ cand,
                                                             const int* __restrict__ cand_count)
{
    __shared__ int lcnt, lbase;
    const int total = *cand_count;
    for (int j0 = (int)(blockIdx.x * 256); j0 < total; j0 += (int)(gridDim.x * 256)) {
        const int j = j0 + (int)threadIdx.x;
        Dists q;
        bool sub = false;
        int64_t gind = 0;
        int g = 0;
        if (j < total) {
            gind = cand[j];
            g = grid_of(B, gind);
            const BGrid G = B.grids[g];
            const Box b = box_of(G);
            int r, c, p;
            coords_of(b, gind - G.off, r, c, p);
            const T* f = F + G.off;
            const double sp[3] = {b.dx, b.dy, b.dz};
#pragma unroll 1
            for (int a = 0; a < 3; ++a) {
                double v[7];
                line7(f, r, c, p, a, b, v);
                q.r[a] = sp[a];
                q.l[a] = sp[a];
                if (v[3] * v[4] < 0)
                    q.r[a] = (ORDER == 2)   ? eno_quadratic_distance(v[2], v[3], v[4], v[5], sp[a])
                             : (ORDER == 3) ? cubic_distance(v[2], v[3], v[4], v[5], sp[a])
                                            : six_distance(v[1], v[2], v[3], v[4], v[5], v[6], sp[a]);
                if (v[2] * v[3] < 0)
                    q.l[a] = sp[a] - ((ORDER == 2)   ? eno_quadratic_distance(v[1], v[2], v[3], v[4], sp[a])
                                      : (ORDER == 3) ? cubic_distance(v[1], v[2], v[3], v[4], sp[a])
                                                     : six_distance(v[0], v[1], v[2], v[3], v[4], v[5], sp[a]));
            }
            // a distance that differs from the spacing (the quadratic ENO
            // locator is not clamped and can land just above it)
            sub = q.r[0] != b.dx || q.l[0] != b.dx || q.r[1] != b.dy || q.l[1] != b.dy || q.r[2] != b.dz ||
                  q.l[2] != b.dz;
            if (sub) flags[gind] |= (uint8_t)FLAG_SUBCELL;
        }
        const int slot = block_slot(sub, il.count, &lcnt, &lbase);
        if (slot >= 0) {
            const Box b = box_of(B.grids[g]);
            il.idx[slot] = (int)gind;
            B.list_grid[slot] = g;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                il.dist[(size_t)(2 * a) * il.cap + slot] = q.r[a];
                il.dist[(size_t)(2 * a + 1) * il.cap + slot] = q.l[a];
            }
            il.dt[slot] = frozen(q, b.dx) ? -1.0 : 0.3 * min6(q);
        }
    }
}

template <typename T>
__global__ void kb_set_mask_bit(uint8_t* __restrict__ flags, const T* __restrict__ F, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t f = flags[i] & (uint8_t)~FLAG_MASK;
    if (F[i] < 0) f |= FLAG_MASK;
    flags[i] = f;
}

// ---------------------------------------------------------------------------
// fused stage kernels
// ---------------------------------------------------------------------------
// Fused stage kernels (one launch per stage, see k_reinit_stage and
// k_extend_stage): the first niface blocks walk the interface list, the
// others take one z-march work item each (a tile column of one grid over a
// range of its pages) with the single-grid kernels' walk and per-point
// bodies, called with the item's grid.
template <int R, typename T, int S = IFACE_WENO>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(stage_waves<T>()))) kb_reinit_stage(T* __restrict__ out, const T* __restrict__ fs,
                                                       const T* __restrict__ fn, const uint8_t* __restrict__ flags,
                                                       IList il, BatchView B, int stage, int* __restrict__ over,
                                                       int* __restrict__ over_count, int seq, int niface)
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
                const int64_t gind = il.idx[al.point];
                const BGrid G = B.grids[B.list_grid[al.point]];
                q = reinit_iface_lane<T, S>(fs + G.off, flags + G.off, box_of(G), inv_of(G), il, al.point,
                                            gind - G.off, gind, al.axis);
            }
            reinit_iface_finish(out, fs, fn, q, al, active, stage);
        }
        return;
    }
    const ZItem it = B.items[xcd_remap(blk - niface, (int)gridDim.x - niface)];
    const BGrid G = B.grids[it.g];
    const Box b = box_of(G);
    const Inv invs = inv_of(G);
    const idx_t off = G.off, plane = (idx_t)b.rows * b.cols;
    T* o = out + off;
    ZFace<T> zf;
    zmarch<R, T, false, true>(lds[0], lds[1], b, it.r0, it.c0, it.p0, it.p1, fs + off, shift(fn, off), nullptr, nullptr,
                        nullptr, flags + off, stage,
                        [&](const T* l, const ZCol<R>& zc, const T w[7], const ZStep<R, T, false>& st, int p) {
                            bool cross;
                            const T res = reinit_value<R>(l, zc.base, w, st.fn, st.flags(zc), invs, stage, cross,
                                                          zf, p == it.p0);
                            const bool bulk = !(st.flags(zc) & FLAG_SUBCELL);
                            if constexpr (defers_crossings<T>()) {
                                if (bulk && !cross) stp(o + (idx_t)p * plane, zc.in, res);
                                wave_append(bulk && cross, (int)(off + (idx_t)p * plane + zc.in), over,
                                            &over_count[seq & 1]);
                            } else if (bulk) {
                                stp(o + (idx_t)p * plane, zc.in, res);
                            }
                        });
}

// launch 2 of 2: the deferred bulk points (see k_reinit_deferred)
template <typename T, int S = IFACE_WENO>
__global__ void __launch_bounds__(256) kb_reinit_deferred(T* __restrict__ out, const T* __restrict__ fs,
                                                          const T* __restrict__ fn, const uint8_t* __restrict__ flags,
                                                          IList il, BatchView B, int stage,
                                                          const int* __restrict__ over,
                                                          const int* __restrict__ over_count, int seq)
{
    const int total = over_count[seq & 1];
    for (int j0 = 0; j0 < total; j0 += axis_lane_stride()) {
        const AxisLane al = axis_lane(j0);
        const bool active = al.point >= 0 && al.point < total;
        ReinitLane q;
        if (active) {
            const int64_t gind = over[al.point];
            const BGrid G = B.grids[grid_of(B, gind)];
            q = reinit_iface_lane<T, S>(fs + G.off, flags + G.off, box_of(G), inv_of(G), il, -1, gind - G.off,
                                        gind, al.axis);
        }
        reinit_iface_finish(out, fs, fn, q, al, active, stage);
    }
}

// fused extension stage (one launch): the first niface blocks walk the
// interface list (three lanes per point), the others take the z-march work
// items; see k_extend_stage
template <int R, typename T, int S = IFACE_WENO>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(stage_waves<T>()))) kb_extend_stage(T* __restrict__ out, const T* __restrict__ cs,
                                                       const T* __restrict__ cn, const T* __restrict__ vx,
                                                       const T* __restrict__ vy, const T* __restrict__ vz,
                                                       IList il, const double* __restrict__ lcp,
                                                       const double* __restrict__ lv, BatchView B, int stage,
                                                       int niface)
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
                const int64_t gind = il.idx[al.point];
                const BGrid G = B.grids[B.list_grid[al.point]];
                q = extend_iface_lane<T, S>(cs + G.off, box_of(G), inv_of(G), il, lcp, lv, al.point, gind - G.off,
                                            gind, al.axis);
            }
            extend_iface_finish(out, cs, cn, q, al, active, stage);
        }
        return;
    }
    const ZItem it = B.items[xcd_remap(blk - niface, (int)gridDim.x - niface)];
    const BGrid G = B.grids[it.g];
    const Box b = box_of(G);
    const Inv inv = inv_of(G);
    const idx_t off = G.off, plane = (idx_t)b.rows * b.cols;
    T* o = out + off;
    // the listed points carry a NaN x-velocity (kb_list_gather_vel)
    zmarch<R, T, true, false>(lds[0], lds[1], b, it.r0, it.c0, it.p0, it.p1, cs + off, shift(cn, off), vx + off,
                       vy + off, vz + off, nullptr, stage,
                       [&](const T* l, const ZCol<R>& zc, const T w[7], const ZStep<R, T, true>& st, int p) {
                           const T res = upwind_value<R>(l, zc.base, w, st.vel, st.fn, (T)inv.dt_bulk, inv, stage);
                           if (zc.valid && st.vel[0] == st.vel[0]) stp(o + (idx_t)p * plane, zc.in, res);
                       });
}

template <int R, typename T>
__global__ void __launch_bounds__(256) kb_advect_stage(T* __restrict__ out, const T* __restrict__ fs,
                                                       const T* __restrict__ fn, const T* __restrict__ vx,
                                                       const T* __restrict__ vy, const T* __restrict__ vz,
                                                       const double* __restrict__ dt, BatchView B, int stage)
{
    __shared__ T lds[2][Shape<R>::N];
    const ZItem it = B.items[xcd_remap((int)blockIdx.x, (int)gridDim.x)];
    const BGrid G = B.grids[it.g];
    const Box b = box_of(G);
    const Inv inv = inv_of(G);
    const T dtg = (T)dt[it.g];
    const idx_t off = G.off, plane = (idx_t)b.rows * b.cols;
    T* o = out + off;
    zmarch<R, T, true, false>(lds[0], lds[1], b, it.r0, it.c0, it.p0, it.p1, fs + off, shift(fn, off), vx + off, vy + off,
                       vz + off, nullptr, stage,
                       [&](const T* l, const ZCol<R>& zc, const T w[7], const ZStep<R, T, true>& st, int p) {
                           const T res = upwind_value<R>(l, zc.base, w, st.vel, st.fn, dtg, inv, stage);
                           if (zc.valid) stp(o + (idx_t)p * plane, zc.in, res);
                       });
}

// ---------------------------------------------------------------------------
// extension velocity: upwind normal (upwind_normal.cu:4-104) rounded to T, then
// v = sign(F) n/max(|n|, floor) (WENO5RK3Extend.m:23-31), as the single-grid
// k_upwind_normal + k_extension_velocity pair computes it.  Points without
// subcell distances in a tile pass, listed points in a list pass.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void store_extension_velocity(T* vx, T* vy, T* vz, const T* F, idx_t i, const double n[3],
                                                         double floor_)
{
    const double fx = (double)(T)n[0], fy = (double)(T)n[1], fz = (double)(T)n[2];
    const double g = dmax(sqrt(fx * fx + fy * fy + fz * fz), floor_);
    const double s = msign((double)F[i]);
    vx[i] = (T)(s * (fx / g));
    vy[i] = (T)(s * (fy / g));
    vz[i] = (T)(s * (fz / g));
}

template <int R, typename T>
__global__ void __launch_bounds__(256) kb_ext_vel_bulk(T* __restrict__ vx, T* __restrict__ vy, T* __restrict__ vz,
                                                       const T* __restrict__ F, const uint8_t* __restrict__ flags,
                                                       double floor_, BatchView B)
{
    // one z-march work item per block (x- and y-taps from the staged tile,
    // z-taps from the register window); only the side |F(+-1)| selects is
    // evaluated, in the Jiang-Peng form of weno5_pair
    __shared__ T lds[2][Shape<R>::N];
    const ZItem it = B.items[xcd_remap((int)blockIdx.x, (int)gridDim.x)];
    const BGrid G = B.grids[it.g];
    const Box b = box_of(G);
    const idx_t off = G.off, plane = (idx_t)b.rows * b.cols;
    const double sp[3] = {b.dx, b.dy, b.dz};
    zmarch<R, T, false, true>(lds[0], lds[1], b, it.r0, it.c0, it.p0, it.p1, F + off, nullptr, nullptr, nullptr,
                              nullptr, flags + off, 0,
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
                                      const JPd j = jpd(sp[a]);
                                      const JPcd<double> m = jp_centre(v);
                                      n[a] = (fabs(v[4]) < fabs(v[2]))
                                                 ? (m.c + jp_q(m.d[4], m.d[3], m.d[2], m.d[1], j.e)) * j.k
                                                 : (m.c - jp_q(m.d[0], m.d[1], m.d[2], m.d[3], j.e)) * j.k;
                                  }
                                  store_extension_velocity(vx, vy, vz, F, (idx_t)(off + (idx_t)p * plane + zc.in), n,
                                                           floor_);
                              });
}

// S: the interface scheme of the listed points' upwind normal: the
// nonuniform WENO (upwind_normal.cu), the cubic ENO where the axis has a
// subcell distance (wenork3_extend.cu:202-292, interface values 0), or ENO2
// (enork2_extend.cu:102-138: the ENO2 derivative on the side whose node
// value, 0 at a crossing, is smaller)
// the ENO-RK2 upwind normal at the points without a subcell distance
template <int R>
__global__ void __launch_bounds__(256) kb_eno2_ext_vel_bulk(double* __restrict__ vx, double* __restrict__ vy,
                                                            double* __restrict__ vz, const double* __restrict__ F,
                                                            const uint8_t* __restrict__ flags, double floor_,
                                                            BatchView B)
{
    const TileRange tr = tile_range(B.ntiles);
    for (int tile = tr.first; tile < tr.end; tile += tr.step) {
        const BGrid G = B.grids[B.tile_grid[tile]];
        const Box b = box_of(G);
        const Tile t = tile_at<R>(b, tile - G.tile0);
        if (!t.valid) continue;
        const idx_t ind = lin(b, t.r, t.c, t.p);
        if (flags[G.off + ind] & FLAG_SUBCELL) continue;
        const double sp[3] = {b.dx, b.dy, b.dz};
        double n[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double v[7], sR, sL;
            line7(F + G.off, t.r, t.c, t.p, a, b, v);
            eno2_pair(sR, sL, v, sp[a], sp[a], sp[a]);
            n[a] = (fabs(v[4]) < fabs(v[2])) ? sR : sL;
        }
        store_extension_velocity(vx, vy, vz, F, (idx_t)(G.off + ind), n, floor_);
    }
}

template <typename T, int S = IFACE_WENO>
__global__ void __launch_bounds__(256) kb_ext_vel_list(T* __restrict__ vx, T* __restrict__ vy, T* __restrict__ vz,
                                                       const T* __restrict__ F, IList il, double floor_, BatchView B)
{
    const int cnt = *il.count;
    for (int j = (int)(blockIdx.x * 256 + threadIdx.x); j < cnt; j += (int)(gridDim.x * 256)) {
        const int64_t gind = il.idx[j];
        const BGrid G = B.grids[B.list_grid[j]];
        const Box b = box_of(G);
        int r, c, p;
        coords_of(b, gind - G.off, r, c, p);
        const double sp[3] = {b.dx, b.dy, b.dz};
        double n[3];
#pragma unroll 1
        for (int a = 0; a < 3; ++a) {
            double v[7], fo, bk;
            line7(F + G.off, r, c, p, a, b, v);
            const double r1 = il.dist[(size_t)(2 * a) * il.cap + j], l1 = il.dist[(size_t)(2 * a + 1) * il.cap + j];
            if constexpr (S == IFACE_ENO2) {
                eno2_pair(fo, bk, v, r1, l1, sp[a]);
                const double vr = (r1 == sp[a]) ? v[4] : 0.0, vl = (l1 == sp[a]) ? v[2] : 0.0;
                n[a] = (fabs(vr) < fabs(vl)) ? fo : bk;
            } else {
                if constexpr (S == IFACE_CUBIC)
                    wenork3_pair(fo, bk, v, r1, l1, sp[a], jpd(sp[a]), r1 < sp[a] || l1 < sp[a], 0.0, 0.0);
                else
                    extend_axis(fo, bk, v, r1, l1, sp[a], jpd(sp[a]), 0.0, 0.0);
                n[a] = (fabs(v[4]) < fabs(v[2])) ? fo : bk;
            }
        }
        store_extension_velocity(vx, vy, vz, F, (idx_t)gind, n, floor_);
    }
}

// interface values at the listed points: the quintic of
// boundary_interpolant.cu:102-141, the cubic of wenork3_extend.cu:310-329 or
// the quadratic ENO of enork2_extend.cu:144-170 (S)
template <typename T, int S = IFACE_WENO>
__global__ void __launch_bounds__(256) kb_list_interp(double* __restrict__ cp, IList il, const T* __restrict__ C,
                                                      BatchView B)
{
    const int total = *il.count;
    for (int j0 = 0; j0 < total; j0 += axis_lane_stride()) {
        const AxisLane al = axis_lane(j0);
        if (!(al.point >= 0 && al.point < total)) continue;
        const int a = al.axis, j = al.point;
        const int64_t gind = il.idx[j];
        const BGrid G = B.grids[B.list_grid[j]];
        const Box b = box_of(G);
        const double sp[3] = {b.dx, b.dy, b.dz};
        const double rf = il.dist[(size_t)(2 * a) * il.cap + j], lb = il.dist[(size_t)(2 * a + 1) * il.cap + j];
        Pair q{0.0, 0.0};
        if constexpr (S == IFACE_ENO2) {
            int r, c, p;
            coords_of(b, gind - G.off, r, c, p);
            double v[7];
            line7(C + G.off, r, c, p, a, b, v);
            eno2_interface_values(q.fore, q.back, rf, lb, sp[a], v);
        } else if (rf < sp[a] || lb < sp[a]) {
            int r, c, p;
            coords_of(b, gind - G.off, r, c, p);
            double v[7];
            line7(C + G.off, r, c, p, a, b, v);
            if constexpr (S == IFACE_CUBIC) cubic_interface_values(q.fore, q.back, rf, lb, sp[a], v);
            else q = sixth_interp(rf, lb, sp[a], v[0], v[1], v[2], v[3], v[4], v[5], v[6]);
        }
        cp[(size_t)(2 * a) * il.cap + j] = q.fore;
        cp[(size_t)(2 * a + 1) * il.cap + j] = q.back;
    }
}

// ENO-RK2 extension stage (ENORK2Extend.m:68-79, Heun; heun2 == 0:
// out = cs - L(cs), heun2 == 1: out = (cn + cs - L(cs)) / 2), as the
// single-grid k_enork2_extend_stage, in one launch: the first niface blocks
// take the listed points (distances, and interface values in list order),
// the others the points without a subcell distance in a z-marching pass over
// the batch's work items (spacings for distances, pseudo-time step 0.3 min of
// them; x- and y-taps from the staged tile, z-taps from the register window,
// the same values line7 reads).  Each point reads cn only at itself, so out
// may be cn.
template <int R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(stage_waves<double>())))
kb_eno2_extend_stage(double* __restrict__ out, const double* __restrict__ cs, const double* __restrict__ cn,
                     const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
                     const uint8_t* __restrict__ flags, IList il, const double* __restrict__ lcp, BatchView B,
                     int heun2, int niface)
{
#pragma clang fp contract(off)
    __shared__ double lds[2][Shape<R>::N];
    if ((int)blockIdx.x < niface) {
        const int cnt = *il.count;
        for (int j = (int)(blockIdx.x * 256 + threadIdx.x); j < cnt; j += niface * 256) {
            const int64_t i = il.idx[j];
            const BGrid G = B.grids[B.list_grid[j]];
            const Box b = box_of(G);
            int r, c, p;
            coords_of(b, i - G.off, r, c, p);
            Dists q;
            double hf[3], hb[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                q.r[a] = il.dist[(size_t)(2 * a) * il.cap + j];
                q.l[a] = il.dist[(size_t)(2 * a + 1) * il.cap + j];
                hf[a] = lcp[(size_t)(2 * a) * il.cap + j];
                hb[a] = lcp[(size_t)(2 * a + 1) * il.cap + j];
            }
            const double step =
                enork2_residual(cs + G.off, r, c, p, b, q, hf, hb, vx[i], vy[i], vz[i], 0.3 * min6(q));
            out[i] = heun2 ? (cn[i] + cs[i] - step) / 2 : cs[i] - step;
        }
        return;
    }
    const ZItem it = B.items[xcd_remap((int)blockIdx.x - niface, (int)gridDim.x - niface)];
    const BGrid G = B.grids[it.g];
    const Box b = box_of(G);
    const idx_t off = G.off, plane = (idx_t)b.rows * b.cols;
    double* o = out + off;
    Dists q;
    q.r[0] = q.l[0] = b.dx;
    q.r[1] = q.l[1] = b.dy;
    q.r[2] = q.l[2] = b.dz;
    const double dt = 0.3 * min6(q);
    // stage != 0 loads the start value cn at each point (ZStep::fn)
    zmarch<R, double, true, true>(lds[0], lds[1], b, it.r0, it.c0, it.p0, it.p1, cs + off, shift(cn, off), vx + off,
                                  vy + off, vz + off, flags + off, heun2,
                                  [&](const double* l, const ZCol<R>& zc, const double w[7],
                                      const ZStep<R, double, true>& st, int p) {
                                      double v[7], fo[3], bk[3];
                                      tile_line_x<R>(l, zc.base, v);
                                      eno2_field_pair(fo[0], bk[0], v, b.dx, b.dx, b.dx, 0.0, 0.0);
                                      tile_line_y(l, zc.base, v);
                                      eno2_field_pair(fo[1], bk[1], v, b.dy, b.dy, b.dy, 0.0, 0.0);
                                      eno2_field_pair(fo[2], bk[2], w, b.dz, b.dz, b.dz, 0.0, 0.0);
                                      const double step = upwind_sum(fo, bk, st.vel[0], st.vel[1], st.vel[2], dt);
                                      const double res = heun2 ? (st.fn + w[3] - step) / 2 : w[3] - step;
                                      if (!(st.flags(zc) & FLAG_SUBCELL)) stp(o + (idx_t)p * plane, zc.in, res);
                                  });
}

template <typename T>
__global__ void __launch_bounds__(256) kb_list_gather_vel(double* __restrict__ lv, IList il, T* __restrict__ vx,
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

// u.N, N = grad F/max(|grad F|,1e-14), clamped central differences
// (set_calculus_toolbox.cu:31-44), as k_normal_velocity
template <int R, typename T>
__global__ void __launch_bounds__(256) kb_normal_velocity(T* __restrict__ out, const T* __restrict__ F,
                                                          const T* __restrict__ vx, const T* __restrict__ vy,
                                                          const T* __restrict__ vz, BatchView B)
{
    const TileRange tr = tile_range(B.ntiles);
    for (int tile = tr.first; tile < tr.end; tile += tr.step) {
        const BGrid G = B.grids[B.tile_grid[tile]];
        const Box b = box_of(G);
        const Tile t = tile_at<R>(b, tile - G.tile0);
        if (!t.valid) continue;
        const int r = t.r, c = t.c, p = t.p;
        const T* f = F + G.off;
        const idx_t ind = G.off + lin(b, r, c, p);
        const double fx = ((double)f[lin(b, r, min(c + 1, b.cols - 1), p)] - (double)f[lin(b, r, max(c - 1, 0), p)]) /
                          (2 * b.dx);
        const double fy = ((double)f[lin(b, min(r + 1, b.rows - 1), c, p)] - (double)f[lin(b, max(r - 1, 0), c, p)]) /
                          (2 * b.dy);
        const double fz = ((double)f[lin(b, r, c, min(p + 1, b.pges - 1))] - (double)f[lin(b, r, c, max(p - 1, 0))]) /
                          (2 * b.dz);
        const double g = dmax(sqrt(fx * fx + fy * fy + fz * fz), 1e-14);
        out[ind] = (T)((double)vx[ind] * (fx / g) + (double)vy[ind] * (fy / g) + (double)vz[ind] * (fz / g));
    }
}

// ---------------------------------------------------------------------------
// per-grid reductions: one compensated partial per tile, then one block per
// grid merges its tiles' partials in tile order (deterministic)
// ---------------------------------------------------------------------------
// cosine-smoothed delta/H of half-width eps (src/0_0_setup/+SD/@SDF3/SDF3.m:126-136),
// |grad F| by periodic central differences (src/0_0_setup/+SD/@GD3/GD3.m:212-223)
template <int R, typename T>
__global__ void __launch_bounds__(256) kb_area_volume_cosine(double* __restrict__ partial, const T* __restrict__ F,
                                                             double eps_factor, BatchView B)
{
    const double pi = 3.14159265358979323846;
    const TileRange tr = tile_range(B.ntiles);
    for (int tile = tr.first; tile < tr.end; tile += tr.step) {
        const BGrid G = B.grids[B.tile_grid[tile]];
        const Box b = box_of(G);
        const Tile t = tile_at<R>(b, tile - G.tile0);
        const double eps = eps_factor * b.dx;
        KSum k[2] = {{0.0, 0.0}, {0.0, 0.0}};
        if (t.valid) {
            const T* f = F + G.off;
            const int r = t.r, c = t.c, p = t.p;
            const double v = (double)f[lin(b, r, c, p)];
            double delta = 0.0, heav;
            if (v > eps) {
                heav = 1.0;
            } else if (v < -eps) {
                heav = 0.0;
            } else {
                const int cp = (c + 1 == b.cols) ? 0 : c + 1, cm = (c == 0) ? b.cols - 1 : c - 1;
                const int rp = (r + 1 == b.rows) ? 0 : r + 1, rm = (r == 0) ? b.rows - 1 : r - 1;
                const int pp = (p + 1 == b.pges) ? 0 : p + 1, pm = (p == 0) ? b.pges - 1 : p - 1;
                const double fx = ((double)f[lin(b, r, cp, p)] - (double)f[lin(b, r, cm, p)]) / (2 * b.dx);
                const double fy = ((double)f[lin(b, rp, c, p)] - (double)f[lin(b, rm, c, p)]) / (2 * b.dy);
                const double fz = ((double)f[lin(b, r, c, pp)] - (double)f[lin(b, r, c, pm)]) / (2 * b.dz);
                const double g = sqrt(fx * fx + fy * fy + fz * fz);
                delta = g * (1 + cos(pi * v / eps)) / eps / 2;
                heav = (1 + v / eps + sin(pi * v / eps) / pi) / 2;
            }
            k[0].add(delta);
            k[1].add(1 - heav);
        }
        block_reduce<2>(k);
        if (threadIdx.x == 0) {
            partial[(size_t)tile * 4 + 0] = k[0].s;
            partial[(size_t)tile * 4 + 1] = k[0].c;
            partial[(size_t)tile * 4 + 2] = k[1].s;
            partial[(size_t)tile * 4 + 3] = k[1].c;
        }
        __syncthreads();
    }
}

// one block per grid: result[2g+q] = (sum of its tiles' partials) * dx*dy*dz
__global__ void __launch_bounds__(256) kb_grid_sum2(double* __restrict__ result, const double* __restrict__ partial,
                                                    BatchView B)
{
    const BGrid G = B.grids[blockIdx.x];
    const int t0 = G.tile0, nt = G.ntiles;
    KSum k[2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int i = (int)threadIdx.x; i < nt; i += 256)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            KSum o{partial[(size_t)(t0 + i) * 4 + 2 * q], partial[(size_t)(t0 + i) * 4 + 2 * q + 1]};
            k[q].merge(o);
        }
    block_reduce<2>(k);
    if (threadIdx.x == 0) {
        const double vol = G.dx * G.dy * G.dz;
        result[2 * blockIdx.x] = (k[0].s + k[0].c) * vol;
        result[2 * blockIdx.x + 1] = (k[1].s + k[1].c) * vol;
    }
}

template <int R, typename T>
__global__ void __launch_bounds__(256) kb_max_abs3(double* __restrict__ partial, const T* __restrict__ vx,
                                                   const T* __restrict__ vy, const T* __restrict__ vz, BatchView B)
{
    __shared__ double red[4];
    const TileRange tr = tile_range(B.ntiles);
    for (int tile = tr.first; tile < tr.end; tile += tr.step) {
        const BGrid G = B.grids[B.tile_grid[tile]];
        const Box b = box_of(G);
        const Tile t = tile_at<R>(b, tile - G.tile0);
        double m = 0.0;
        if (t.valid) {
            const idx_t i = G.off + lin(b, t.r, t.c, t.p);
            m = dmax(fabs((double)vx[i]), dmax(fabs((double)vy[i]), fabs((double)vz[i])));
        }
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) m = dmax(m, __shfl_xor(m, s, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) partial[tile] = dmax(dmax(red[0], red[1]), dmax(red[2], red[3]));
        __syncthreads();
    }
}

// one block per grid: dt[g] = cfl*dx/max|u| (the CFL step of
// p00_test_singlephase_relax.m:89, in that operation order)
__global__ void __launch_bounds__(256) kb_grid_dt(double* __restrict__ dt, const double* __restrict__ partial,
                                                  double cfl, BatchView B)
{
    __shared__ double red[4];
    const BGrid G = B.grids[blockIdx.x];
    const int t0 = G.tile0, nt = G.ntiles;
    double m = 0.0;
    for (int i = (int)threadIdx.x; i < nt; i += 256) m = dmax(m, partial[t0 + i]);
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) m = dmax(m, __shfl_xor(m, s, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) dt[blockIdx.x] = cfl * G.dx / dmax(dmax(red[0], red[1]), dmax(red[2], red[3]));
}

// ---------------------------------------------------------------------------
// synthetic inputs
// ---------------------------------------------------------------------------
template <int R, typename T>
__global__ void __launch_bounds__(256) kb_curl_velocity(T* __restrict__ vx, T* __restrict__ vy, T* __restrict__ vz,
                                                        const double* __restrict__ modes, int nmodes,
                                                        const double* __restrict__ lo, BatchView B)
{
    const TileRange tr = tile_range(B.ntiles);
    for (int tile = tr.first; tile < tr.end; tile += tr.step) {
        const int g = B.tile_grid[tile];
        const BGrid G = B.grids[g];
        const Box b = box_of(G);
        const Tile t = tile_at<R>(b, tile - G.tile0);
        if (!t.valid) continue;
        const double x = lo[3 * g] + t.c * b.dx, y = lo[3 * g + 1] + t.r * b.dy, z = lo[3 * g + 2] + t.p * b.dz;
        double u = 0.0, v = 0.0, w = 0.0;
        const double* md = modes + (size_t)g * nmodes * 7;
        for (int m = 0; m < nmodes; ++m, md += 7) {
            const double s = -sin(md[0] * x + md[1] * y + md[2] * z + md[6]);
            u += s * md[3];
            v += s * md[4];
            w += s * md[5];
        }
        const idx_t i = G.off + lin(b, t.r, t.c, t.p);
        vx[i] = (T)u;
        vy[i] = (T)v;
        vz[i] = (T)w;
    }
}

template <int R, typename T>
__global__ void __launch_bounds__(256) kb_ellipsoid(T* __restrict__ F, const double* __restrict__ axes,
                                                    const double* __restrict__ lo, BatchView B)
{
    const TileRange tr = tile_range(B.ntiles);
    for (int tile = tr.first; tile < tr.end; tile += tr.step) {
        const int g = B.tile_grid[tile];
        const BGrid G = B.grids[g];
        const Box b = box_of(G);
        const Tile t = tile_at<R>(b, tile - G.tile0);
        if (!t.valid) continue;
        const double x = lo[3 * g] + t.c * b.dx, y = lo[3 * g + 1] + t.r * b.dy, z = lo[3 * g + 2] + t.p * b.dz;
        const double a = axes[3 * g], bb = axes[3 * g + 1], c = axes[3 * g + 2];
        F[G.off + lin(b, t.r, t.c, t.p)] = (T)(sqrt(x * x / (a * a) + y * y / (bb * bb) + z * z / (c * c)) - 1.0);
    }
}

// ===========================================================================
// host launchers
// ===========================================================================
// a tile-loop kernel with the batch's tile shape (B and s in scope)
#define TR_LAUNCH(KERNEL, T, GRID, ...)                                 \
    do {                                                                \
        if (B.tr == 32) KERNEL<32, T><<<GRID, 256, 0, s>>>(__VA_ARGS__); \
        else KERNEL<64, T><<<GRID, 256, 0, s>>>(__VA_ARGS__);           \
    } while (0)

static inline int list_blocks(const BatchView& B)
{
    // one sweep of an interface-list launch covers 84 points per block
    const int b = B.ntiles / 12;
    return b < 16 ? 16 : (b > 4096 ? 4096 : b);
}

// interface-list blocks of a fused stage launch: two per CU at most (see
// fused_iface_blocks of the single-grid launches)
static int fused_list_blocks(const BatchView& B)
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
    const int b = list_blocks(B);
    return b < 2 * cus ? b : 2 * cus;
}

static inline IList ilist(const ListView& lv) { return IList{lv.idx, lv.count, lv.dist, lv.dt, lv.cap}; }

template <typename T>
hipError_t launch_batch_build_list(const BatchView& B, const T* F, const T* mask_src, uint8_t* flags,
                                   const ListView& lv, int order, int* cand, int* cand_count, hipStream_t s)
{
    TR_LAUNCH(kb_crossings, T, B.blocks, F, mask_src, flags, B, cand, cand_count);
    const int nb = list_blocks(B);
    if (order == 2) kb_crossing_distances<T, 2><<<nb, 256, 0, s>>>(F, flags, lv, B, cand, cand_count);
    else if (order == 3) kb_crossing_distances<T, 3><<<nb, 256, 0, s>>>(F, flags, lv, B, cand, cand_count);
    else kb_crossing_distances<T, 6><<<nb, 256, 0, s>>>(F, flags, lv, B, cand, cand_count);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_set_mask_bit(uint8_t* flags, const T* F, int64_t n, hipStream_t s)
{
    kb_set_mask_bit<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(flags, F, (idx_t)n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_reinit_stage(const BatchView& B, T* out, const T* fs, const T* fn, const uint8_t* flags,
                                     const ListView& lv, int stage, int* over, int* over_count, int seq,
                                     hipStream_t s, int scheme)
{
    const int ni = fused_list_blocks(B);
    const int nd = ni / 10 < 8 ? 8 : ni / 10;
    if constexpr (std::is_same<T, double>::value) {
        if (scheme == IFACE_CUBIC) {
            if (B.tr == 32)
                kb_reinit_stage<32, T, IFACE_CUBIC><<<B.nitems + ni, 256, 0, s>>>(out, fs, fn, flags, ilist(lv), B,
                                                                                  stage, over, over_count, seq, ni);
            else
                kb_reinit_stage<64, T, IFACE_CUBIC><<<B.nitems + ni, 256, 0, s>>>(out, fs, fn, flags, ilist(lv), B,
                                                                                  stage, over, over_count, seq, ni);
            kb_reinit_deferred<T, IFACE_CUBIC><<<nd, 256, 0, s>>>(out, fs, fn, flags, ilist(lv), B, stage,
                                                                    (const int*)over, (const int*)over_count, seq);
            return hipGetLastError();
        }
    }
    if (scheme != IFACE_WENO) return hipErrorInvalidValue;
    TR_LAUNCH(kb_reinit_stage, T, B.nitems + ni, out, fs, fn, flags, ilist(lv), B, stage, over, over_count, seq, ni);
    if constexpr (defers_crossings<T>())
        kb_reinit_deferred<T><<<nd, 256, 0, s>>>(out, fs, fn, flags, ilist(lv), B, stage, (const int*)over,
                                                   (const int*)over_count, seq);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_extend_stage(const BatchView& B, T* out, const T* cs, const T* cn, const T* vx,
                                     const T* vy, const T* vz, const ListView& lv, const double* lcp,
                                     const double* lvel, int stage, hipStream_t s, int scheme)
{
    const int ni = fused_list_blocks(B);
    if constexpr (std::is_same<T, double>::value) {
        if (scheme == IFACE_CUBIC) {
            if (B.tr == 32)
                kb_extend_stage<32, T, IFACE_CUBIC><<<B.nitems + ni, 256, 0, s>>>(out, cs, cn, vx, vy, vz,
                                                                                  ilist(lv), lcp, lvel, B, stage, ni);
            else
                kb_extend_stage<64, T, IFACE_CUBIC><<<B.nitems + ni, 256, 0, s>>>(out, cs, cn, vx, vy, vz,
                                                                                  ilist(lv), lcp, lvel, B, stage, ni);
            return hipGetLastError();
        }
    }
    if (scheme != IFACE_WENO) return hipErrorInvalidValue;
    TR_LAUNCH(kb_extend_stage, T, B.nitems + ni, out, cs, cn, vx, vy, vz, ilist(lv), lcp, lvel, B, stage, ni);
    return hipGetLastError();
}

hipError_t launch_batch_eno2_extend_stage(const BatchView& B, double* out, const double* cs, const double* cn,
                                          const double* vx, const double* vy, const double* vz, const uint8_t* flags,
                                          const ListView& lv, const double* lcp, int heun2, hipStream_t s)
{
    const int ni = fused_list_blocks(B);
    if (B.tr == 32)
        kb_eno2_extend_stage<32><<<B.nitems + ni, 256, 0, s>>>(out, cs, cn, vx, vy, vz, flags, ilist(lv), lcp, B,
                                                               heun2, ni);
    else
        kb_eno2_extend_stage<64><<<B.nitems + ni, 256, 0, s>>>(out, cs, cn, vx, vy, vz, flags, ilist(lv), lcp, B,
                                                               heun2, ni);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_advect_stage(const BatchView& B, T* out, const T* fs, const T* fn, const T* vx,
                                     const T* vy, const T* vz, const double* dt, int stage, hipStream_t s)
{
    TR_LAUNCH(kb_advect_stage, T, B.nitems, out, fs, fn, vx, vy, vz, dt, B, stage);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_extension_velocity(const BatchView& B, T* vx, T* vy, T* vz, const T* F,
                                           const uint8_t* flags, const ListView& lv, double floor_, hipStream_t s,
                                           int scheme)
{
    if constexpr (std::is_same<T, double>::value) {
        if (scheme == IFACE_ENO2) {
            if (B.tr == 32) kb_eno2_ext_vel_bulk<32><<<B.blocks, 256, 0, s>>>(vx, vy, vz, F, flags, floor_, B);
            else kb_eno2_ext_vel_bulk<64><<<B.blocks, 256, 0, s>>>(vx, vy, vz, F, flags, floor_, B);
            kb_ext_vel_list<T, IFACE_ENO2><<<list_blocks(B), 256, 0, s>>>(vx, vy, vz, F, ilist(lv), floor_, B);
            return hipGetLastError();
        }
        if (scheme == IFACE_CUBIC) {
            TR_LAUNCH(kb_ext_vel_bulk, T, B.nitems, vx, vy, vz, F, flags, floor_, B);
            kb_ext_vel_list<T, IFACE_CUBIC><<<list_blocks(B), 256, 0, s>>>(vx, vy, vz, F, ilist(lv), floor_, B);
            return hipGetLastError();
        }
    }
    if (scheme != IFACE_WENO) return hipErrorInvalidValue;
    TR_LAUNCH(kb_ext_vel_bulk, T, B.nitems, vx, vy, vz, F, flags, floor_, B);
    kb_ext_vel_list<T><<<list_blocks(B), 256, 0, s>>>(vx, vy, vz, F, ilist(lv), floor_, B);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_list_interp(const BatchView& B, double* lcp, const ListView& lv, const T* C, hipStream_t s,
                                    int scheme)
{
    if constexpr (std::is_same<T, double>::value) {
        if (scheme == IFACE_CUBIC) {
            kb_list_interp<T, IFACE_CUBIC><<<list_blocks(B), 256, 0, s>>>(lcp, ilist(lv), C, B);
            return hipGetLastError();
        }
        if (scheme == IFACE_ENO2) {
            kb_list_interp<T, IFACE_ENO2><<<list_blocks(B), 256, 0, s>>>(lcp, ilist(lv), C, B);
            return hipGetLastError();
        }
    }
    if (scheme != IFACE_WENO) return hipErrorInvalidValue;
    kb_list_interp<T><<<list_blocks(B), 256, 0, s>>>(lcp, ilist(lv), C, B);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_list_gather_vel(double* lvel, const ListView& lv, T* vx, const T* vy, const T* vz,
                                        hipStream_t s)
{
    kb_list_gather_vel<T><<<1024, 256, 0, s>>>(lvel, ilist(lv), vx, vy, vz);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_normal_velocity(const BatchView& B, T* out, const T* F, const T* vx, const T* vy,
                                        const T* vz, hipStream_t s)
{
    TR_LAUNCH(kb_normal_velocity, T, B.blocks, out, F, vx, vy, vz, B);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_area_volume_cosine(const BatchView& B, double* result, double* partial, const T* F,
                                           double eps_factor, hipStream_t s)
{
    TR_LAUNCH(kb_area_volume_cosine, T, B.blocks, partial, F, eps_factor, B);
    kb_grid_sum2<<<B.ngrids, 256, 0, s>>>(result, partial, B);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_cfl_dt(const BatchView& B, double* dt, double* partial, const T* vx, const T* vy,
                               const T* vz, double cfl, hipStream_t s)
{
    TR_LAUNCH(kb_max_abs3, T, B.blocks, partial, vx, vy, vz, B);
    kb_grid_dt<<<B.ngrids, 256, 0, s>>>(dt, partial, cfl, B);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_curl_velocity(const BatchView& B, T* vx, T* vy, T* vz, const double* modes, int nmodes,
                                      const double* lo, hipStream_t s)
{
    TR_LAUNCH(kb_curl_velocity, T, B.blocks, vx, vy, vz, modes, nmodes, lo, B);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_ellipsoid(const BatchView& B, T* F, const double* axes, const double* lo, hipStream_t s)
{
    TR_LAUNCH(kb_ellipsoid, T, B.blocks, F, axes, lo, B);
    return hipGetLastError();
}

#define VLS_BATCH_INSTANTIATE(T)                                                                                    \
    template hipError_t launch_batch_build_list<T>(const BatchView&, const T*, const T*, uint8_t*, const ListView&, \
                                                   int, int*, int*, hipStream_t);                                   \
    template hipError_t launch_batch_set_mask_bit<T>(uint8_t*, const T*, int64_t, hipStream_t);                     \
    template hipError_t launch_batch_reinit_stage<T>(const BatchView&, T*, const T*, const T*, const uint8_t*,      \
                                                     const ListView&, int, int*, int*, int, hipStream_t, int);      \
    template hipError_t launch_batch_extend_stage<T>(const BatchView&, T*, const T*, const T*, const T*, const T*,  \
                                                     const T*, const ListView&, const double*, const double*, int,  \
                                                     hipStream_t, int);                                             \
    template hipError_t launch_batch_advect_stage<T>(const BatchView&, T*, const T*, const T*, const T*, const T*,  \
                                                     const T*, const double*, int, hipStream_t);                    \
    template hipError_t launch_batch_extension_velocity<T>(const BatchView&, T*, T*, T*, const T*, const uint8_t*,  \
                                                           const ListView&, double, hipStream_t, int);              \
    template hipError_t launch_batch_list_interp<T>(const BatchView&, double*, const ListView&, const T*,           \
                                                    hipStream_t, int);                                              \
    template hipError_t launch_batch_list_gather_vel<T>(double*, const ListView&, T*, const T*, const T*,           \
                                                        hipStream_t);                                               \
    template hipError_t launch_batch_normal_velocity<T>(const BatchView&, T*, const T*, const T*, const T*,         \
                                                        const T*, hipStream_t);                                     \
    template hipError_t launch_batch_area_volume_cosine<T>(const BatchView&, double*, double*, const T*, double,    \
                                                           hipStream_t);                                            \
    template hipError_t launch_batch_cfl_dt<T>(const BatchView&, double*, double*, const T*, const T*, const T*,    \
                                               double, hipStream_t);                                                \
    template hipError_t launch_batch_curl_velocity<T>(const BatchView&, T*, T*, T*, const double*, int,             \
                                                      const double*, hipStream_t);                                  \
    template hipError_t launch_batch_ellipsoid<T>(const BatchView&, T*, const double*, const double*, hipStream_t);

VLS_BATCH_INSTANTIATE(double)
VLS_BATCH_INSTANTIATE(float)
#undef VLS_BATCH_INSTANTIATE

}  // namespace vls
