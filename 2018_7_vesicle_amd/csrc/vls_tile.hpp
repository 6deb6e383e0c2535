// Device pieces shared by the single-grid kernels (vls_kernels.hip) and the
// batched kernels (vls_batch.hip): launch geometry, RK combinations, interface
// lists, plane tiles, axis-lane grouping and compensated block reductions.
#pragma once
#include "vls_device.hpp"
#include "vls_kernels.hpp"

namespace vls {

// ---------------------------------------------------------------------------
// launch geometry helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool point_of_thread(const Box& b, int& r, int& c, int& p)
{
    r = blockIdx.x * blockDim.x + threadIdx.x;
    c = blockIdx.y * blockDim.y + threadIdx.y;
    p = blockIdx.z;
    return r < b.rows && c < b.cols;
}

__device__ __forceinline__ idx_t lin(const Box& b, int r, int c, int p)
{
    return (idx_t)p * b.rows * b.cols + (idx_t)c * b.rows + r;
}

// 1/dx, 1/dy, 1/dz and the pseudo-time step of a point without subcell
// distances, 0.3*min(dx,dx,dy,dy,dz,dz) (setDistance.m:17-21 with every
// distance equal to the spacing), computed once on the host
struct Inv {
    double i[3];
    double dt_bulk;
    JPd jd[3];  // per-axis constants of the WENO5 form (set_jp)
    JPf jf[3];
};

// the per-axis constants of the WENO5 helpers for FP64 and FP32 fields
__device__ __forceinline__ JPd axis_arg(const Inv& v, int a, double) { return v.jd[a]; }
__device__ __forceinline__ JPf axis_arg(const Inv& v, int a, float) { return v.jf[a]; }

// RK combinations in the order the MATLAB wrappers evaluate them
// (+SD/@SDF3/WENO5RK3Reinitialization.m:20-37), without fused multiply-adds.
__device__ __forceinline__ double rk_combine(int stage, double fs, double step, double fn)
{
#pragma clang fp contract(off)
    const double t = fs - step;
    if (stage == 0) return t;
    if (stage == 1) return 3. / 4. * fn + 1. / 4. * t;
    return 1. / 3. * fn + 2. / 3. * t;
}

__device__ __forceinline__ float rk_combine(int stage, float fs, float step, float fn)
{
    const float t = fs - step;
    if (stage == 0) return t;
    if (stage == 1) return fmaf(0.75f, fn, 0.25f * t);
    return fmaf(1.f / 3.f, fn, (2.f / 3.f) * t);
}

// the interface test of re_step.cu:260-267 at the middle node of v[2..4]:
// a sign change between it and either neighbour, the reference's product test
__device__ __forceinline__ bool sign_change3(const double v[7]) { return (v[2] * v[3] < 0) | (v[3] * v[4] < 0); }

// append `take` lanes' values to list[*count ...] with one atomic per wave
__device__ __forceinline__ void wave_append(bool take, int value, int* __restrict__ list, int* __restrict__ count)
{
    const unsigned long long m = __ballot(take);
    if (m == 0) return;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader, 64);
    if (take) list[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}

// Block-aggregated list append: the lanes of a block collect their entries
// in LDS and the block moves them to the list with one global atomic per
// flush.  A wave-aggregated append costs one global atomic per wave, and
// when many waves append, those atomics serialize on the one counter.  Every
// thread of the block calls push() the same number of times and flush() once
// at the end; both contain block barriers.
template <int CAP>
struct BlockList {
    int* buf;    // LDS, CAP entries
    int* count;  // LDS
    int* base;   // LDS
    __device__ __forceinline__ void init()
    {
        if (threadIdx.x == 0) *count = 0;
        __syncthreads();
    }
    __device__ __forceinline__ void flush(int* __restrict__ list, int* __restrict__ list_count)
    {
        __syncthreads();
        const int n = *count;  // the same in every thread after the barrier
        if (n > 0) {
            if (threadIdx.x == 0) *base = atomicAdd(list_count, n);
            __syncthreads();
            const int b = *base;
            for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) list[b + i] = buf[i];
            __syncthreads();
            if (threadIdx.x == 0) *count = 0;
            __syncthreads();
        }
    }
    __device__ __forceinline__ void push(bool take, int value, int* __restrict__ list, int* __restrict__ list_count)
    {
        __syncthreads();  // the previous push's entries are in, so *count is final
        if (*count > CAP - (int)blockDim.x) flush(list, list_count);
        const unsigned long long m = __ballot(take);
        if (m == 0) return;
        const int lane = (int)__lane_id();
        const int leader = __ffsll((long long)m) - 1;
        int off = 0;
        if (lane == leader) off = atomicAdd(count, __popcll(m));
        off = __shfl(off, leader, 64);
        if (take) buf[off + __popcll(m & ((1ull << lane) - 1ull))] = value;
    }
};

// Block-aggregated slot assignment: every thread of the block calls it; the
// taking lanes get consecutive slots of the list counted by *count, with one
// global atomic for the whole block (lds_cnt, lds_base: two LDS ints).
__device__ __forceinline__ int block_slot(bool take, int* __restrict__ count, int* lds_cnt, int* lds_base)
{
    __syncthreads();  // the previous call's readers of *lds_base are done
    if (threadIdx.x == 0) *lds_cnt = 0;
    __syncthreads();
    const unsigned long long m = __ballot(take);
    const int lane = (int)__lane_id();
    int off = 0;
    if (m != 0) {
        const int leader = __ffsll((long long)m) - 1;
        if (lane == leader) off = atomicAdd(lds_cnt, __popcll(m));
        off = __shfl(off, leader, 64) + __popcll(m & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if (threadIdx.x == 0) *lds_base = (*lds_cnt > 0) ? atomicAdd(count, *lds_cnt) : 0;
    __syncthreads();
    return take ? *lds_base + off : -1;
}

// as wave_append, returning each taking lane's slot (-1 for the others)
__device__ __forceinline__ int wave_append_slot(bool take, int* __restrict__ count)
{
    const unsigned long long m = __ballot(take);
    if (m == 0) return -1;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader, 64);
    return take ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}

struct Dists {
    double r[3], l[3];
};

// an interface list in the form the stage kernels read: indices and, in list
// order, the six distances (row k of `dist` is xpr,xpl,ypf,ypb,zpu,zpd for
// k = 0..5, rows `cap` apart) and the pseudo-time step 0.3*min_dist, or -1 for
// a frozen point
struct IList {
    const int* idx;
    const int* count;
    const double* dist;
    const double* dt;
    int cap;
};
struct IListW {
    const int* idx;
    const int* count;
    double* dist;
    double* dt;
    int cap;
};

__device__ __forceinline__ bool frozen(const Dists& q, double dx)
{
    const double eps = 1e-6 * dx;
    return q.r[0] < eps || q.l[0] < eps || q.r[1] < eps || q.l[1] < eps || q.r[2] < eps || q.l[2] < eps;
}

__device__ __forceinline__ double min6(const Dists& q)
{
    return dmin(dmin(q.r[0], q.l[0]), dmin(dmin(q.r[1], q.l[1]), dmin(q.r[2], q.l[2])));
}

// per-axis derivative pair with the sign-change interface test (re_step.cu:248-269);
// INL: the nonuniform face inline (see weno_nonuniform_pair_inl)
// Interface schemes of the fused stage kernels (IFACE_WENO, IFACE_CUBIC of
// vls_kernels.hpp): the WENO5 path's nonuniform WENO (re_step.cu:79-243,
// shared_utilities.cup:49-278) or the WENO-RK3 variants' cubic ENO
// (weno_rk3_reinitialization.cu:149-247, wenork3_extend.cu:88-190).  Away from
// the interface both use the same uniform WENO5, so the bulk of a stage is
// shared.

// the upwind combination of extend_step.cu:121-123 (enork2_extend.cu:262-264,
// wenork3_extend.cu:490-492), summed left to right, times dt
__device__ __forceinline__ double upwind_sum(const double fo[3], const double bk[3], double vx, double vy, double vz,
                                             double dt)
{
#pragma clang fp contract(off)
    return (dmin(0.0, vx) * fo[0] + dmax(0.0, vx) * bk[0] + dmin(0.0, vy) * fo[1] + dmax(0.0, vy) * bk[1] +
            dmin(0.0, vz) * fo[2] + dmax(0.0, vz) * bk[2]) *
           dt;
}

// ENO-RK2 extension step at a point (enork2_extend.cu:213-266; no frozen
// test) with the interface values of each axis given (hf, hb: fore and back;
// each is used only where that side's distance differs from ds)
__device__ __forceinline__ double enork2_residual(const double* __restrict__ f, int r, int c, int p, const Box& b,
                                                  const Dists& q, const double hf[3], const double hb[3], double vx,
                                                  double vy, double vz, double dt)
{
    const double sp[3] = {b.dx, b.dy, b.dz};
    double fo[3], bk[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v[7];
        line7(f, r, c, p, a, b, v);
        eno2_field_pair(fo[a], bk[a], v, q.r[a], q.l[a], sp[a], (q.r[a] != sp[a]) ? hf[a] : 0.0,
                        (q.l[a] != sp[a]) ? hb[a] : 0.0);
    }
    return upwind_sum(fo, bk, vx, vy, vz, dt);
}

// per-axis pair with the sign-change interface test in the cubic-ENO scheme
__device__ __forceinline__ void reinit_axis_cubic(double& fore, double& back, const double v[7], double r1,
                                                  double l1, double ds, JPd jp)
{
    wenork3_pair(fore, back, v, r1, l1, ds, jp, v[2] * v[3] < 0 || v[3] * v[4] < 0, 0.0, 0.0);
}

template <bool INL = false>
__device__ __forceinline__ void reinit_axis(double& fore, double& back, const double v[7], double r1, double l1,
                                            double ds, JPd jp)
{
    if (v[2] * v[3] < 0 || v[3] * v[4] < 0) {
        Pair q;
        if constexpr (INL) q = weno_nonuniform_pair_inl(v[0], v[1], v[2], v[3], v[4], v[5], v[6], r1, l1, ds, 0.0, 0.0);
        else q = weno_nonuniform_pair(v, r1, l1, ds, 0.0, 0.0);
        fore = q.fore;
        back = q.back;
    } else {
        weno5_pair(fore, back, v, jp);
    }
}

// (row, col, page) of element ind of a grid; 32-bit division (a grid holds
// fewer than 2^31 points: interface lists store int indices)
__device__ __forceinline__ void coords_of(const Box& b, idx_t ind, int& r, int& c, int& p)
{
    const unsigned u = (unsigned)ind, rows = (unsigned)b.rows, plane = rows * (unsigned)b.cols;
    const unsigned pp = u / plane, rem = u - pp * plane, cc = rem / rows;
    p = (int)pp;
    c = (int)cc;
    r = (int)(rem - cc * rows);
}

__device__ __forceinline__ void load_dists(Dists& q, const Dist6c& d, idx_t ind)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) { q.r[a] = d.r[a][ind]; q.l[a] = d.l[a][ind]; }
}

// ---------------------------------------------------------------------------
// plane tiles: a block covers 64 x 4 (rows x cols) points of one page.  The
// page's values with a 3-point halo in rows and cols are staged in LDS once
// (clamped there), so the x- and y-taps of the WENO stencil are LDS reads at
// constant offsets; the z-taps are global loads at a wave-uniform page offset.
// ---------------------------------------------------------------------------
constexpr int HALO = 3;

// tile shapes: R rows x 256/R cols of one page (R = 64: a wave is one column
// of 64 rows; R = 32: a wave is two columns of 32 rows, for boxes whose
// rows are a multiple of 32 but not of 64), staged with a 3-point halo
template <int R>
struct Shape {
    static constexpr int TR = R, TC = 256 / R, LR = R + 2 * HALO, LC = TC + 2 * HALO, N = LR * LC;
};

struct Tile {
    int r0, c0, p;  // tile origin (row, col) and page
    int r, c;       // this thread's point
    bool valid;
    int base;       // LDS index of this thread's point
    idx_t page;     // element offset of page p in the box (uniform)
    unsigned in;    // element offset of the point in its page, c * rows + r
};

// element `off` of a page: the uniform 64-bit page base plus a 32-bit byte
// offset per lane, so a load or store is one instruction with a scalar base
// (offsets reach 7 pages of doubles: pages hold fewer than 2^26 points,
// vls_sdf3_create and vls_batch_create refuse larger ones)
template <typename T>
__device__ __forceinline__ T ldp(const T* __restrict__ base, unsigned off)
{
    return *(const T*)((const char*)base + off * (unsigned)sizeof(T));
}
// as ldp, marked non-temporal: for a stream read once per launch (the
// velocity of the extension and advection stages), so that its lines do not
// push the staged tiles' halos out of L2 (measured: extension stage at 256^3
// FP32 103-107 -> 98-100 us)
template <typename T>
__device__ __forceinline__ T ldp_nt(const T* __restrict__ base, unsigned off)
{
    return __builtin_nontemporal_load((const T*)((const char*)base + off * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void stp(T* __restrict__ base, unsigned off, T v)
{
    *(T*)((char*)base + off * (unsigned)sizeof(T)) = v;
}

// Blocks are dispatched round-robin over the 8 XCDs (block i to XCD i % 8),
// each with its own L2.  Give every XCD a contiguous range of tiles instead,
// so neighbouring tiles share their halos in one L2.
__device__ __forceinline__ int xcd_remap(int blk, int nblk)
{
    if ((nblk & 7) != 0) return blk;
    return (blk & 7) * (nblk >> 3) + (blk >> 3);
}

// tiles per column and per row of a page, and in a box
template <int R>
__host__ __device__ inline int tiles_per_col(int rows) { return (rows + R - 1) / R; }
template <int R>
__host__ __device__ inline int tiles_per_row(int cols) { return (cols + Shape<R>::TC - 1) / Shape<R>::TC; }
template <int R>
__host__ __device__ inline int box_tiles(int rows, int cols, int pges)
{
    return tiles_per_col<R>(rows) * tiles_per_row<R>(cols) * pges;
}

// tile number `blk` of the box (rows fastest, then cols, then pages)
template <int R>
__device__ __forceinline__ Tile tile_at(const Box& b, int blk)
{
    using S = Shape<R>;
    const int ntr = tiles_per_col<R>(b.rows), ntc = tiles_per_row<R>(b.cols);
    Tile t;
    const int rb = blk % ntr;
    const int q = blk / ntr;
    const int cb = q % ntc;
    t.p = q / ntc;
    t.r0 = rb * S::TR;
    t.c0 = cb * S::TC;
    const int ty = (int)(threadIdx.x % S::TR), tx = (int)(threadIdx.x / S::TR);
    t.r = t.r0 + ty;
    t.c = t.c0 + tx;
    t.valid = t.r < b.rows && t.c < b.cols;
    t.base = (tx + HALO) * S::LR + ty + HALO;
    t.page = (idx_t)t.p * b.rows * b.cols;
    t.in = (unsigned)t.c * (unsigned)b.rows + (unsigned)t.r;
    return t;
}

// the tiles a block of a persistent launch visits: XCD x (= blockIdx.x % 8,
// the hardware's round-robin) takes the contiguous range
// [x*chunk, (x+1)*chunk) and its blocks stride through it
struct TileRange {
    int first, end, step;
};

__device__ __forceinline__ TileRange tile_range(int ntiles)
{
    TileRange r;
    const int nb = (int)gridDim.x;
    if ((nb & 7) == 0) {
        const int x = (int)(blockIdx.x & 7), j = (int)(blockIdx.x >> 3), per = nb >> 3;
        const int chunk = (ntiles + 7) >> 3;
        r.first = x * chunk + j;
        r.end = min(ntiles, (x + 1) * chunk);
        r.step = per;
    } else {
        r.first = (int)blockIdx.x;
        r.end = ntiles;
        r.step = nb;
    }
    return r;
}

// A staged tile: one page's values of a tile with its clamped 3-point halo
// in lds[Shape<R>::N], column-major (LR rows per column).  The loads go to
// registers first (zstep_load) and store_tile writes them to LDS, so a
// kernel issues every other load of its points in between and waits for
// memory once.  Each thread loads one row of the first TR staged rows in
// every TC-th staged column (the column offsets are wave-uniform for R = 64),
// and threads below 6*LC one element of the last 6 staged rows.
template <int R, typename T>
struct TileRegs {
    static constexpr int NK = (Shape<R>::LC + Shape<R>::TC - 1) / Shape<R>::TC;
    T m[NK];
    T s;
};

template <int R>
__device__ __forceinline__ int tile_tx()
{
    const int tx = (int)threadIdx.x / R;
    return (R % 64 == 0) ? __builtin_amdgcn_readfirstlane(tx) : tx;
}

template <int R, typename T>
__device__ __forceinline__ void store_tile(T* lds, const TileRegs<R, T>& g)
{
    using S = Shape<R>;
    const int ty = (int)threadIdx.x % R, tx = tile_tx<R>();
#pragma unroll
    for (int k = 0; k < TileRegs<R, T>::NK; ++k) {
        const int lc = tx + S::TC * k;
        if (lc < S::LC) lds[lc * S::LR + ty] = g.m[k];
    }
    const int e = (int)threadIdx.x;
    if (e < 2 * HALO * S::LC) lds[(e / (2 * HALO)) * S::LR + S::TR + e % (2 * HALO)] = g.s;
}

template <typename T>
__device__ __forceinline__ void tile_line_y(const T* lds, int base, T v[7])
{
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] = lds[base + k - 3];
}

template <int R, typename T>
__device__ __forceinline__ void tile_line_x(const T* lds, int base, T v[7])
{
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] = lds[base + (k - 3) * Shape<R>::LR];
}

// Interface work is spread over three lanes per point, one per axis (21
// points per wave64, lane 63 idle): the per-axis chains are long (the
// nonuniform WENO divides ~30 times) and the lists short, so this triples the
// parallelism.  The axis results meet in the group's first lane, which adds
// them in the reference's x, y, z order.
struct AxisLane {
    int point;  // index into the point list, or -1
    int axis;   // 0 x, 1 y, 2 z
    int base;   // lane of the group's x member
};

// the lane's group in sweep j0 of block blk (of a 256-thread launch)
__device__ __forceinline__ AxisLane axis_lane(int j0, int blk)
{
    const int lane = (int)(threadIdx.x & 63);
    const int wave = blk * 4 + (int)(threadIdx.x >> 6);
    AxisLane a;
    a.axis = lane % 3;
    a.base = lane - a.axis;
    a.point = (lane < 63) ? j0 + wave * 21 + lane / 3 : -1;
    return a;
}
__device__ __forceinline__ AxisLane axis_lane(int j0) { return axis_lane(j0, (int)blockIdx.x); }

// stride, in points, between the groups nblk blocks visit per sweep
__device__ __forceinline__ int axis_lane_stride(int nblk) { return nblk * 4 * 21; }
__device__ __forceinline__ int axis_lane_stride() { return axis_lane_stride((int)gridDim.x); }

// per-axis pair with the distance interface test (shared_utilities.cup:256-278)
__device__ __forceinline__ void extend_axis(double& fore, double& back, const double v[7], double r1, double l1,
                                            double ds, JPd jp, double h_fore, double h_back)
{
    if (r1 < ds || l1 < ds) {
        const Pair q = weno_nonuniform_pair(v, r1, l1, ds, h_fore, h_back);
        fore = q.fore;
        back = q.back;
    } else {
        weno5_pair(fore, back, v, jp);
    }
}

// Neumaier accumulator
struct KSum {
    double s, c;
    __device__ __forceinline__ void add(double x)
    {
        const double t = s + x;
        c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
        s = t;
    }
    __device__ __forceinline__ void merge(const KSum& o)
    {
        add(o.s);
        c += o.c;
    }
};

__device__ __forceinline__ double shfl_xor_d(double v, int m)
{
    return __shfl_xor(v, m, 64);
}

__device__ __forceinline__ void wave_reduce(KSum& k)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        KSum o;
        o.s = shfl_xor_d(k.s, m);
        o.c = shfl_xor_d(k.c, m);
        k.merge(o);
    }
}

// block of 256 threads -> thread 0 holds the merged (s,c) of each quantity
template <int NQ>
__device__ __forceinline__ void block_reduce(KSum (&k)[NQ])
{
    __shared__ double lds[NQ][2][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        wave_reduce(k[q]);
        if (lane == 0) { lds[q][0][wave] = k[q].s; lds[q][1][wave] = k[q].c; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            KSum t{0.0, 0.0};
            for (int w = 0; w < 4; ++w) { KSum o{lds[q][0][w], lds[q][1][w]}; t.merge(o); }
            k[q] = t;
        }
    }
}

template <int NQ>
__device__ __forceinline__ void block_reduce_store(KSum (&k)[NQ], double* __restrict__ partial)
{
    block_reduce<NQ>(k);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            partial[(size_t)blockIdx.x * 2 * NQ + 2 * q] = k[q].s;
            partial[(size_t)blockIdx.x * 2 * NQ + 2 * q + 1] = k[q].c;
        }
}


// ---------------------------------------------------------------------------
// per-point bodies of the fused stage kernels.  The single-grid kernels and
// the batched kernels call the same functions, with the field pointers of one
// grid (offset into the packed fields for a batch), so a grid gives the same
// result bit for bit either way.
// ---------------------------------------------------------------------------

// Waves per SIMD the fused stage kernels are compiled for: their bulk
// z-march fits 128 VGPRs in FP64 and 96 in FP32; the interface lanes (the
// nonuniform WENO, inlined) are fitted into the same budget.
template <typename T>
constexpr int stage_waves()
{
    return sizeof(T) == 8 ? 4 : 5;
}

// ---------------------------------------------------------------------------
// z-marching walk (single grid): a block takes one tile column (TR rows x TC
// cols of every page) over pages [p0, p1).  The z-taps of a point are a
// register window over pages p-3..p+3 that shifts by one page per step, so a
// step loads one new z-value per point; the tile and the point's other inputs
// of page p+1 are loaded while page p is evaluated (two LDS buffers, one
// barrier per page).  In-page offsets are computed once per block; moving to
// the next page only moves the (scalar) page base.
// ---------------------------------------------------------------------------
template <int R>
struct ZCol {
    unsigned m[TileRegs<R, double>::NK];  // in-page offsets of the tile-row loads
    unsigned s;                           // in-page offset of the strip load
    unsigned in;                          // the point's in-page offset (0 outside the box)
    int base;                             // the point's LDS index
    bool valid;                           // the point is in the box
};

template <int R>
__device__ __forceinline__ ZCol<R> zcol(const Box& b, int r0, int c0)
{
    using S = Shape<R>;
    ZCol<R> z;
    const int ty = (int)threadIdx.x % R, tx = tile_tx<R>();
    const unsigned rows = (unsigned)b.rows;
    const unsigned row = (unsigned)clampi(r0 - HALO + ty, b.rows - 1);
#pragma unroll
    for (int k = 0; k < TileRegs<R, double>::NK; ++k) {
        const int lc = tx + S::TC * k;
        z.m[k] = (lc < S::LC) ? (unsigned)clampi(c0 - HALO + lc, b.cols - 1) * rows + row : 0u;
    }
    const int e = (int)threadIdx.x;
    z.s = (e < 2 * HALO * S::LC) ? (unsigned)clampi(c0 - HALO + e / (2 * HALO), b.cols - 1) * rows +
                                       (unsigned)clampi(r0 - HALO + S::TR + e % (2 * HALO), b.rows - 1)
                                 : 0u;
    const int r = r0 + ty, c = c0 + tx;
    z.valid = r < b.rows && c < b.cols;
    z.in = z.valid ? (unsigned)c * rows + (unsigned)r : 0u;
    z.base = (tx + HALO) * S::LR + ty + HALO;
    return z;
}

// the loads of one page: the staged tile, the point's velocity, start value
// and flags, and the z-value the window takes in at this page.  The loaded
// values are kept as loaded; nothing is computed from them until the page is
// evaluated, so the loads of page p+1 stay in flight while page p is (an
// operation on them here would make the compiler wait for every load first).
template <int R, typename T, bool VEL>
struct ZStep {
    TileRegs<R, T> reg;
    T vel[VEL ? 3 : 1];
    T fn, w;  // fn: the start value (at stage 0 the point's own value, unused)
    uint8_t fl;  // the point's flags as loaded
    // the point's flags, FLAG_SUBCELL for a lane outside the box (never written)
    __device__ __forceinline__ uint8_t flags(const ZCol<R>& zc) const { return zc.valid ? fl : (uint8_t)FLAG_SUBCELL; }
};

// Every load is issued by every lane on every path: the lanes without a tile
// element to stage load element 0 of the page (their offsets are 0), and the
// march always loads a next page (the last page again at its end).  The
// compiler's wait counts then do not depend on the path taken, so a page's
// evaluation never waits for the loads of the page after it.  FL: the point
// flags are loaded (else 0).
template <int R, typename T, bool VEL, bool FL>
__device__ __forceinline__ void zstep_load(ZStep<R, T, VEL>& s, const ZCol<R>& zc, int p, const Box& b, idx_t plane,
                                           const T* __restrict__ fs, const T* __restrict__ fn,
                                           const T* __restrict__ vx, const T* __restrict__ vy,
                                           const T* __restrict__ vz, const uint8_t* __restrict__ flags, int stage)
{
    const idx_t po = (idx_t)p * plane;
    const T* fp = fs + po;
#pragma unroll
    for (int k = 0; k < TileRegs<R, T>::NK; ++k) s.reg.m[k] = ldp(fp, zc.m[k]);
    s.reg.s = ldp(fp, zc.s);
    s.w = ldp(fs + (idx_t)min(p + 3, b.pges - 1) * plane, zc.in);
    if (VEL) {
        s.vel[0] = ldp_nt(vx + po, zc.in);
        s.vel[1] = ldp_nt(vy + po, zc.in);
        s.vel[2] = ldp_nt(vz + po, zc.in);
    }
    // at stage 0 (no start value) the point's own value is read and not used
    // (rk_combine does not read fn at stage 0)
    s.fn = ldp((stage != 0 ? fn : fs) + po, zc.in);
    if (FL) s.fl = ldp(flags + po, zc.in);
    else s.fl = 0;
}

template <typename T>
__device__ __forceinline__ void window_shift(T w[7], T in)
{
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = w[k + 1];
    w[6] = in;
}

// body(lds, zc, w, step, p) evaluates page p: x- and y-taps from lds, z-taps w
template <int R, typename T, bool VEL, bool FL, class Body>
__device__ __forceinline__ void zmarch(T* lds0, T* lds1, const Box& b, int r0, int c0, int p0, int p1,
                                       const T* __restrict__ fs, const T* __restrict__ fn,
                                       const T* __restrict__ vx, const T* __restrict__ vy,
                                       const T* __restrict__ vz, const uint8_t* __restrict__ flags, int stage,
                                       Body body)
{
    if (p0 >= p1) return;
    const ZCol<R> zc = zcol<R>(b, r0, c0);
    const idx_t plane = (idx_t)b.rows * b.cols;
    // the window of page p0 - 1 (pages p0-4 .. p0+2, clamped); its first
    // value leaves the window before any page reads it, so it is not loaded
    T w[7];
    w[0] = T(0);
#pragma unroll
    for (int k = 1; k < 7; ++k) w[k] = ldp(fs + (idx_t)clampi(p0 - 4 + k, b.pges - 1) * plane, zc.in);
    ZStep<R, T, VEL> sa, sb;
    zstep_load<R, T, VEL, FL>(sa, zc, p0, b, plane, fs, fn, vx, vy, vz, flags, stage);
    for (int p = p0;;) {
        window_shift(w, sa.w);
        store_tile<R>(lds0, sa.reg);
        __syncthreads();
        zstep_load<R, T, VEL, FL>(sb, zc, min(p + 1, p1 - 1), b, plane, fs, fn, vx, vy, vz, flags, stage);
        body(lds0, zc, w, sa, p);
        if (++p >= p1) break;
        window_shift(w, sb.w);
        store_tile<R>(lds1, sb.reg);
        __syncthreads();
        zstep_load<R, T, VEL, FL>(sa, zc, min(p + 1, p1 - 1), b, plane, fs, fn, vx, vy, vz, flags, stage);
        body(lds1, zc, w, sb, p);
        if (++p >= p1) break;
    }
}

// the z-march share of block blk of the nblk = zmarch_blocks(...) blocks of a
// launch: tile column (r0, c0) and pages [p0, p1).  Consecutive blocks of one
// XCD take neighbouring columns, which share their halos in its L2.
template <int R>
__device__ __forceinline__ void zmarch_share(const Box& b, int depth, int blk, int nblk, int& r0, int& c0, int& p0,
                                             int& p1)
{
    const int ntr = tiles_per_col<R>(b.rows), nct = ntr * tiles_per_row<R>(b.cols);
    const int gi = xcd_remap(blk, nblk);
    const int ch = gi / nct, ct = gi - ch * nct;
    const int cb = ct / ntr;
    r0 = (ct - cb * ntr) * R;
    c0 = cb * Shape<R>::TC;
    p0 = ch * depth;
    p1 = min(p0 + depth, b.pges);
}

// blocks of a z-march launch with `depth` pages per block
template <int R>
__host__ __device__ inline int zmarch_blocks(int rows, int cols, int pges, int depth)
{
    return tiles_per_col<R>(rows) * tiles_per_row<R>(cols) * ((pges + depth - 1) / depth);
}

// Whether the reinitialization bulk defers a point next to a sign change of
// the stage field to the interface path.  The reference evaluates such a
// point with its nonuniform WENO at full spacings (re_step.cu:248-269: the
// point has no subcell distance), which is the uniform WENO5 up to rounding
// (the oracle tests hold the two within 1e-14 relative); FP64 defers so the
// results stay the reference's bit for bit, FP32 evaluates them in its bulk.
template <typename T>
constexpr bool defers_crossings()
{
    return sizeof(T) == sizeof(double);
}

// WENO5-RK3 reinitialization stage at a bulk point: x- and y-taps from the
// staged tile (lds, the point at index base), z-taps z[0..6] (z[3] the point),
// the start value fn and flags fl.  Returns the new value; where the bulk
// defers crossings (FP64), cross is set for a point next to a sign change of
// the stage field (it is left to the interface pass), else it stays false.
// zf: the z-faces the march's previous page left (first: this is the
// block's first page, zf is set up here)
template <int R, typename T>
__device__ __forceinline__ T reinit_value(const T* lds, int base, const T z[7], T fn, uint8_t fl, const Inv& invs,
                                          int stage, bool& cross, ZFace<T>& zf, bool first)
{
    cross = false;
    T fo[3], bk[3];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        T v[7];
        if (a == 0) tile_line_x<R>(lds, base, v);
        else tile_line_y(lds, base, v);
        if constexpr (defers_crossings<T>()) cross |= sign_change3(v);
        weno5_pair(fo[a], bk[a], v, axis_arg(invs, a, T()));
    }
    if (first) zface_start(zf, z, axis_arg(invs, 2, T()));
    if constexpr (defers_crossings<T>()) {
        bool cz;
        weno5_pair_z(fo[2], bk[2], cz, z, zf, axis_arg(invs, 2, T()));
        cross |= cz;
    } else {
        weno5_pair_z(fo[2], bk[2], z, zf, axis_arg(invs, 2, T()));
    }
    const T step = godunov_step((fl & FLAG_MASK) != 0, fo[0], bk[0], fo[1], bk[1], fo[2], bk[2], (T)invs.dt_bulk);
    return rk_combine(stage, z[3], step, fn);
}

// one lane (point, axis) of the reinitialization interface pass
struct ReinitLane {
    double m = 0.0;   // this axis's Godunov term
    double dt = 0.0;  // pseudo-time step; < 0: frozen point
    idx_t gind = 0;   // where the point's result goes
    bool inside = false;
};

// slot: the point's position in the interface list, or -1 for a deferred
// bulk point (spacing distances, bulk time step); ind: its index in the grid
// (fs, flags of that grid); gind: its index in the output fields
template <typename T, int SCHEME = IFACE_WENO>
__device__ __forceinline__ ReinitLane reinit_iface_lane(const T* __restrict__ fs, const uint8_t* __restrict__ flags,
                                                        const Box& b, const Inv& inv, const IList& il, int slot,
                                                        idx_t ind, idx_t gind, int a)
{
    const double sp[3] = {b.dx, b.dy, b.dz};
    ReinitLane q;
    q.gind = gind;
    q.dt = inv.dt_bulk;
    int r, c, p;
    coords_of(b, ind, r, c, p);
    q.inside = (flags[ind] & FLAG_MASK) != 0;
    double r1 = sp[a], l1 = sp[a];
    if (slot >= 0) {
        q.dt = il.dt[slot];
        r1 = il.dist[(size_t)(2 * a) * il.cap + slot];
        l1 = il.dist[(size_t)(2 * a + 1) * il.cap + slot];
    }
    if (q.dt >= 0) {
        double v[7], fo, bk;
        line7(fs, r, c, p, a, b, v);
        if constexpr (SCHEME == IFACE_CUBIC) reinit_axis_cubic(fo, bk, v, r1, l1, sp[a], inv.jd[a]);
        else reinit_axis<true>(fo, bk, v, r1, l1, sp[a], inv.jd[a]);
        q.m = godunov_axis(q.inside, bk, fo);
    }
    return q;
}

// the group's x lane adds the three axis terms in x, y, z order and writes
// the RK combination (every lane of the wave must call this)
template <typename T>
__device__ __forceinline__ void reinit_iface_finish(T* __restrict__ out, const T* __restrict__ fs,
                                                    const T* __restrict__ fn, const ReinitLane& q,
                                                    const AxisLane& al, bool active, int stage)
{
    const double my = __shfl(q.m, al.base + 1, 64), mz = __shfl(q.m, al.base + 2, 64);
    if (active && al.axis == 0) {
        double step = 0.0;  // frozen point (dt < 0)
        if (q.dt >= 0) {
            // godunov_step: (sqrt((mx + my) + mz) - 1) * dt, negated inside
            const double s1 = (sqrt((q.m + my) + mz) - 1) * q.dt;
            step = q.inside ? -s1 : s1;
        }
        out[q.gind] = (T)rk_combine(stage, (double)fs[q.gind], step, (stage == 0) ? 0.0 : (double)fn[q.gind]);
    }
}

// upwind WENO5 rate of the extension and advection stages at a bulk point
// (x- and y-taps from the staged tile, z-taps z[0..6]) and the RK
// combination with the stage's time step: acc * scale is the step
template <int R, typename T>
__device__ __forceinline__ T upwind_value(const T* lds, int base, const T z[7], const T vel[3], T fn, T scale,
                                          const Inv& inv, int stage)
{
    T acc = 0;
    T v[7];
    tile_line_x<R>(lds, base, v);
    acc = upwind_axis(acc, vel[0], v, axis_arg(inv, 0, T()));
    tile_line_y(lds, base, v);
    acc = upwind_axis(acc, vel[1], v, axis_arg(inv, 1, T()));
    acc = upwind_axis(acc, vel[2], z, axis_arg(inv, 2, T()));
    return rk_combine(stage, z[3], acc * scale, fn);
}

// one lane (point, axis) of the extension interface pass
struct ExtendLane {
    double t = 0.0;    // this axis's term coef * derivative
    double dt = -1.0;  // pseudo-time step; < 0: frozen point
    idx_t gind = 0;
};

// slot: position in the list (lcp, lv, il.dist, il.dt rows); ind: index in
// the grid of cs; gind: index in the output fields
template <typename T, int SCHEME = IFACE_WENO>
__device__ __forceinline__ ExtendLane extend_iface_lane(const T* __restrict__ cs, const Box& b, const Inv& inv,
                                                        const IList& il, const double* __restrict__ lcp,
                                                        const double* __restrict__ lv, int slot, idx_t ind,
                                                        idx_t gind, int a)
{
    const double sp[3] = {b.dx, b.dy, b.dz};
    ExtendLane q;
    q.gind = gind;
    q.dt = il.dt[slot];
    if (q.dt >= 0) {
        int r, c, p;
        coords_of(b, ind, r, c, p);
        double v[7];
        line7(cs, r, c, p, a, b, v);
        const double vel = lv[(size_t)a * il.cap + slot];
        const double r1 = il.dist[(size_t)(2 * a) * il.cap + slot];
        const double l1 = il.dist[(size_t)(2 * a + 1) * il.cap + slot];
        if (r1 < sp[a] || l1 < sp[a]) {
            const bool fore = vel < 0;
            const double hf = lcp[(size_t)(2 * a) * il.cap + slot], hb = lcp[(size_t)(2 * a + 1) * il.cap + slot];
            double dv;
            if constexpr (SCHEME == IFACE_CUBIC) {
                double x[7], h[7], df, db;
                cubic_stencil(x, h, v, r1, l1, sp[a], hf, hb);
                eno_cubic_pair(df, db, x, h);
                dv = fore ? df : db;
            } else {
                dv = weno_nonuniform_one_inl(v[0], v[1], v[2], v[3], v[4], v[5], v[6], r1, l1, sp[a], hf, hb, !fore);
            }
            q.t = 0.0 + (fore ? vel : dmax(0.0, vel)) * dv;
        } else {
            q.t = upwind_axis(0.0, vel, v, inv.jd[a]);
        }
    }
    return q;
}

template <typename T>
__device__ __forceinline__ void extend_iface_finish(T* __restrict__ out, const T* __restrict__ cs,
                                                    const T* __restrict__ cn, const ExtendLane& q,
                                                    const AxisLane& al, bool active, int stage)
{
    const double ty = __shfl(q.t, al.base + 1, 64), tz = __shfl(q.t, al.base + 2, 64);
    if (active && al.axis == 0) {
        const double step = (q.dt >= 0) ? ((q.t + ty) + tz) * q.dt : 0.0;
        out[q.gind] = (T)rk_combine(stage, (double)cs[q.gind], step, (stage == 0) ? 0.0 : (double)cn[q.gind]);
    }
}

}  // namespace vls
