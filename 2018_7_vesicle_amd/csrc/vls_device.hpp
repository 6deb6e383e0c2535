// Device arithmetic of the level-set hot path, written for gfx950 (wave64,
// full-rate FP64 VALU).  Every function states the reference kernel it
// reproduces.  The formulas are the reference's; the uniform WENO5 weights use
// one division instead of six.  Everything is compiled without contraction
// (-ffp-contract=off) and every fused multiply-add is written out, so each
// result is one fixed IEEE-754 sequence: the CPU oracle evaluates the same
// sequences and the two agree bit for bit.  Against the reference's own build
// (which contracts freely and calls pow) they agree to rounding.
// Paths are relative to src/7_0_paper/2_0_Vesicle/CUDA_Code/.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vls_kernels.hpp"

namespace vls {

typedef int64_t idx_t;

__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double sq(double a) { return a * a; }
// value part of a scalar (the formulas below are templates on the scalar
// type: double, or a dual number of the adjoint; branch tests read values)
__device__ __forceinline__ double val(double a) { return a; }
// kernel sign(): +1 for x>0, -1 otherwise (1_3_WENO5RK3_Reinitialization/shared_utilities.hpp:17-18)
__device__ __forceinline__ double ksign(double x) { return (x > 0) ? 1.0 : -1.0; }
// MATLAB sign(): 0 at 0 (used by the extension velocity, +SD/@SDF3/WENO5RK3Extend.m:29-31)
__device__ __forceinline__ double msign(double x) { return (x > 0) ? 1.0 : ((x < 0) ? -1.0 : 0.0); }
__device__ __forceinline__ double min_mod(double x, double y)
{
    return (x * y < 0) ? 0.0 : (fabs(x) < fabs(y) ? x : y);
}

// Box geometry: rows = y (contiguous), cols = x, pges = z; clamped
// neighbours (shared_utilities.hpp:21-31).
struct Box {
    int rows, cols, pges;
    double dx, dy, dz;
};

__device__ __forceinline__ int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// the 7 values f(offset -3..+3) along `axis` (0 = x/cols, 1 = y/rows, 2 = z/pges)
template <typename T>
__device__ __forceinline__ void line7(const T* __restrict__ f, int r, int c, int p, int axis,
                                      const Box& b, double out[7])
{
    const idx_t plane = (idx_t)b.rows * b.cols;
    if (axis == 1) {
        const idx_t base = (idx_t)p * plane + (idx_t)c * b.rows;
#pragma unroll
        for (int s = -3; s <= 3; ++s) out[s + 3] = (double)f[base + clampi(r + s, b.rows - 1)];
    } else if (axis == 0) {
        const idx_t base = (idx_t)p * plane + r;
#pragma unroll
        for (int s = -3; s <= 3; ++s) out[s + 3] = (double)f[base + (idx_t)clampi(c + s, b.cols - 1) * b.rows];
    } else {
        const idx_t base = (idx_t)c * b.rows + r;
#pragma unroll
        for (int s = -3; s <= 3; ++s) out[s + 3] = (double)f[base + (idx_t)clampi(p + s, b.pges - 1) * plane];
    }
}

// ---------------------------------------------------------------------------
// Uniform WENO5 derivatives of re_step.cu:11-37 (same text in
// 2_2_WENO5RK3_Extend/shared_utilities.cup:19-45, advection_step.cu:10-28,
// nonoscillating_interpolant.cu:31-57) in the Jiang-Peng form (G.-S. Jiang,
// D. Peng, SIAM J. Sci. Comput. 21 (2000), eq. 2.5-2.8), which is
// algebraically identical to the reference's stencil-weight form and cheaper
// to evaluate.  With nodal values p[0..6] (p[3] the point), undivided second
// differences D_j = p_j - 2 p_{j+1} + p_{j+2} and the fourth-order central
// part C = p_1 - p_5 + 8 (p_4 - p_2):
//   D-  = (C - Q(D0, D1, D2, D3)) / (12 ds),   D+ = (C + Q(D4, D3, D2, D1)) / (12 ds),
//   Q(a,b,c,d) = [4 w0 (a - 2b + c) + (2 w2 - W)(b - 2c + d)] / W,
//   w0 = s1^2 s2^2, w1 = 6 s0^2 s2^2, w2 = 3 s0^2 s1^2, W = w0 + w1 + w2,
//   s0 = (a-b)^2 + 3/13 (a-3b)^2 + e, s1 = (b-c)^2 + 3/13 (b+c)^2 + e,
//   s2 = (c-d)^2 + 3/13 (3c-d)^2 + e,  e = 12/13 * 1e-6 * ds^2,
// where s_k = 12 ds^2/13 * (S_k + 1e-6) of the reference's smoothness
// indicators S_k of divided differences, so the weights are the reference's.
// The FP64 form below is the exact IEEE sequence of the CPU oracle
// (or_weno_pair; the library is compiled without floating-point contraction
// and the fused multiply-adds are explicit); templates on the scalar S also
// evaluate it on dual numbers for the adjoint.  Nothing overflows while the
// second differences stay below about 1e70 (W ~ D^8).
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ S jp_q(S a, S b, S c, S d, double e)
{
    const double k = 3.0 / 13.0;
    const S t0 = a - b, t1 = b - c, t2 = c - d;
    const S y0 = fma(-3.0, b, a), y1 = b + c, y2 = fma(3.0, c, -d);
    const S s0 = fma(t0, t0, fma(k * y0, y0, S(e)));
    const S s1 = fma(t1, t1, fma(k * y1, y1, S(e)));
    const S s2 = fma(t2, t2, fma(k * y2, y2, S(e)));
    const S a0 = s0 * s0, a1 = s1 * s1, a2 = s2 * s2;
    const S w0 = a1 * a2, w1 = 6.0 * (a0 * a2), w2 = 3.0 * (a0 * a1);
    const S W = (w0 + w1) + w2;
    return fma(4.0 * w0, t0 - t1, fma(2.0, w2, -W) * (t1 - t2)) / W;
}

// the second differences D0..D4 and the central part C of p[0..6]
template <typename S>
struct JPcd {
    S d[5];
    S c;
};
template <typename S>
__device__ __forceinline__ JPcd<S> jp_centre(const S p[7])
{
    JPcd<S> m;
#pragma unroll
    for (int j = 0; j < 5; ++j) m.d[j] = fma(-2.0, p[j + 1], p[j]) + p[j + 2];
    m.c = fma(8.0, p[4] - p[2], p[1] - p[5]);
    return m;
}

// Upwind term of one axis, min(0,vel)*D+ + max(0,vel)*D- (extend_step.cu:121-123,
// advection_step.cu:51-53), added to acc in that order.  Only one of the two
// one-sided derivatives has a nonzero weight, so only that one is evaluated:
// the stencil is picked per lane (no divergence) and the other term, an exact
// zero times a finite value, is the +-0 the sum would have added unchanged.
// acc is never -0 (it starts at +0 and x + -x rounds to +0), so the result
// is bit-identical to evaluating both derivatives.
template <typename S>
__device__ __forceinline__ S upwind_axis(S acc, S vel, const S p[7], JPd j)
{
    const JPcd<S> m = jp_centre(p);
    const bool fore = val(vel) < 0;
    const S q = jp_q(fore ? m.d[4] : m.d[0], fore ? m.d[3] : m.d[1], m.d[2], fore ? m.d[1] : m.d[3], j.e);
    const S d = (fore ? m.c + q : m.c - q) * j.k;
    return acc + (fore ? vel : dmax(S(0.0), vel)) * d;
}

// back/fore pair (re_step.cu:260-267)
template <typename S>
__device__ __forceinline__ void weno5_pair(S& d_fore, S& d_back, const S p[7], JPd j)
{
    const JPcd<S> m = jp_centre(p);
    d_back = (m.c - jp_q(m.d[0], m.d[1], m.d[2], m.d[3], j.e)) * j.k;
    d_fore = (m.c + jp_q(m.d[4], m.d[3], m.d[2], m.d[1], j.e)) * j.k;
}

// ---------------------------------------------------------------------------
// FP32 bulk arithmetic, for the points away from the interface: the same
// Jiang-Peng form in single precision, with the weights written through the
// reciprocals of the indicators: w0 : w1 : w2 = s1^2 s2^2 : 6 s0^2 s2^2 :
// 3 s0^2 s1^2 = 1/s0^2 : 6/s1^2 : 3/s2^2 (divide by s0^2 s1^2 s2^2).  Every
// s_k >= e > 0, so the reciprocals are finite; their squares stay in FP32
// range while the indicators lie in [1e-19, 1e19] (second differences up to
// ~1e9), and below that the largest weights still dominate: a weight that
// underflows is one that would not count.  The divisions are hardware
// reciprocals (v_rcp_f32, 1 ulp); the bulk is VALU-bound, and this form needs
// fewer instructions than scaling the indicators by a power of two and
// forming their fourth powers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fmaxf_(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float fminf_(float a, float b) { return (a < b) ? a : b; }

// the correction Q(a, b, c, d) of one side.  The indicators enter the weights
// only through their ratios, so they are taken divided by 1 + 3/13 (the floor
// with them: the FP32 e of the axis constants is e * 13/16) and completed to
// squares:
//   s0 = (a - 11/8 b)^2 + 39/64 b^2 + e',  s1 = (b - 5/8 c)^2 + 39/64 c^2 + e',
//   s2 = (d - 11/8 c)^2 + 39/64 c^2 + e'
// (s1 and s2 share 39/64 c^2 + e').  With x_k = 1/s_k^2 the weights are
// x0 : 6 x1 : 3 x2, and with A = a - 2b + c, B = b - 2c + d the correction is
//   Q = [4 x0 A + (6 x2 - W) B] / W = (4 x0 A + 6 x2 B) / W - B,  W = x0 + 6 x1 + 3 x2.
// W is floored at 1e-30 so that the quotient stays finite; W comes near that
// floor only when every indicator exceeds ~1e15 (second differences beyond
// ~1e7, far outside level-set data).
__device__ __forceinline__ float jp_q(float a, float b, float c, float d, float e)
{
    const float ib = fmaf(0.609375f * b, b, e);
    const float ic = fmaf(0.609375f * c, c, e);
    const float u0 = fmaf(-1.375f, b, a), u1 = fmaf(-0.625f, c, b), u2 = fmaf(-1.375f, c, d);
    const float r0 = __builtin_amdgcn_rcpf(fmaf(u0, u0, ib));
    const float r1 = __builtin_amdgcn_rcpf(fmaf(u1, u1, ic));
    const float r2 = __builtin_amdgcn_rcpf(fmaf(u2, u2, ic));
    const float x0 = r0 * r0, x2 = r2 * r2;
    const float W = fmaf(6.f * r1, r1, fmaf(3.f, x2, x0));
    const float A = fmaf(-2.f, b, a) + c, B = fmaf(-2.f, c, b) + d;
    const float P = fmaf(4.f * x0, A, (6.f * x2) * B);
    return fmaf(P, __builtin_amdgcn_rcpf(__builtin_fmaxf(W, 1e-30f)), -B);
}

// second differences D1, D2, D3 and the central part C of p[0..6]
struct JPc {
    float d1, d2, d3, c;
};
__device__ __forceinline__ JPc jp_centre(const float p[7])
{
    JPc r;
    r.d1 = fmaf(-2.f, p[2], p[1]) + p[3];
    r.d2 = fmaf(-2.f, p[3], p[2]) + p[4];
    r.d3 = fmaf(-2.f, p[4], p[3]) + p[5];
    r.c = fmaf(8.f, p[4] - p[2], p[1] - p[5]);
    return r;
}

// Upwind term of one axis added to acc: vel * (vel < 0 ? D+ : D-), as
// extend_step.cu:121-123 / advection_step.cu:51-53 (min(0,vel) D+ + max(0,vel) D-).
// With kv = vel/(12 ds): vel * D-+ = kv C - |kv| Q on the upwind side.
__device__ __forceinline__ float upwind_axis(float acc, float vel, const float p[7], JPf j)
{
    const JPc m = jp_centre(p);
    const bool fore = vel < 0;
    const float d0 = fmaf(-2.f, p[1], p[0]) + p[2], d4 = fmaf(-2.f, p[5], p[4]) + p[6];
    const float q = jp_q(fore ? d4 : d0, fore ? m.d3 : m.d1, m.d2, fore ? m.d1 : m.d3, j.e);
    const float kv = vel * j.k;
    return fmaf(-__builtin_fabsf(kv), q, fmaf(kv, m.c, acc));
}

// back/fore pair (re_step.cu:260-267).  The two corrections Q(D0,D1,D2,D3)
// and Q(D4,D3,D2,D1) share the term 39/64 D2^2 + e' of their last two
// indicators and B = D1 - 2 D2 + D3; each is jp_q's expression.
__device__ __forceinline__ void weno5_pair(float& d_fore, float& d_back, const float p[7], JPf j)
{
    const JPc m = jp_centre(p);
    const float d0 = fmaf(-2.f, p[1], p[0]) + p[2], d4 = fmaf(-2.f, p[5], p[4]) + p[6];
    const float d1 = m.d1, d2 = m.d2, d3 = m.d3, e = j.e;
    const float ic = fmaf(0.609375f * d2, d2, e);
    const float B = fmaf(-2.f, d2, d1) + d3;
    // back: a = d0, b = d1, c = d2, d = d3
    float qb, qf;
    {
        const float ib = fmaf(0.609375f * d1, d1, e);
        const float u0 = fmaf(-1.375f, d1, d0), u1 = fmaf(-0.625f, d2, d1), u2 = fmaf(-1.375f, d2, d3);
        const float r0 = __builtin_amdgcn_rcpf(fmaf(u0, u0, ib));
        const float r1 = __builtin_amdgcn_rcpf(fmaf(u1, u1, ic));
        const float r2 = __builtin_amdgcn_rcpf(fmaf(u2, u2, ic));
        const float x0 = r0 * r0, x2 = r2 * r2;
        const float W = fmaf(6.f * r1, r1, fmaf(3.f, x2, x0));
        const float A = fmaf(-2.f, d1, d0) + d2;
        const float P = fmaf(4.f * x0, A, (6.f * x2) * B);
        qb = fmaf(P, __builtin_amdgcn_rcpf(__builtin_fmaxf(W, 1e-30f)), -B);
    }
    // fore: a = d4, b = d3, c = d2, d = d1 (its B is the same D3 - 2 D2 + D1)
    {
        const float ib = fmaf(0.609375f * d3, d3, e);
        const float u0 = fmaf(-1.375f, d3, d4), u1 = fmaf(-0.625f, d2, d3), u2 = fmaf(-1.375f, d2, d1);
        const float r0 = __builtin_amdgcn_rcpf(fmaf(u0, u0, ib));
        const float r1 = __builtin_amdgcn_rcpf(fmaf(u1, u1, ic));
        const float r2 = __builtin_amdgcn_rcpf(fmaf(u2, u2, ic));
        const float x0 = r0 * r0, x2 = r2 * r2;
        const float W = fmaf(6.f * r1, r1, fmaf(3.f, x2, x0));
        const float A = fmaf(-2.f, d3, d4) + d2;
        const float P = fmaf(4.f * x0, A, (6.f * x2) * B);
        qf = fmaf(P, __builtin_amdgcn_rcpf(__builtin_fmaxf(W, 1e-30f)), -B);
    }
    d_back = (m.c - qb) * j.k;
    d_fore = (m.c + qf) * j.k;
}

// ---------------------------------------------------------------------------
// z-axis WENO5 pairs along a z-march (one point per page, pages in order).
// The back side of page p+1 and the fore side of page p take the same four
// second differences in reverse order (E_{p-1}..E_{p+2}), so their outer
// smoothness indicators coincide exactly (s0 of one is s2 of the other, the
// same operations on the same operands) and so do the products of the FP64
// weights; the back side of page p+1 takes them from page p.  The results
// are bit for bit those of weno5_pair on the page's 7-point z-line.  ZFace
// holds what page p leaves for page p+1; zface_start sets it up for the
// block's first page from that page's z-line (the fore side of the page
// before, whose result is not used).
// ---------------------------------------------------------------------------
template <typename T>
struct ZFace;

template <>
struct ZFace<float> {
    float d[4];    // D0..D3 of the next page (D1..D4 of this one)
    float ic;      // 39/64 D2^2 + e' of the next page (this page's fore-side ib)
    float x0, x2;  // back-side 1/s0^2, 1/s2^2 of the next page (this page's fore 1/s2^2, 1/s0^2)
    float A;       // back-side A of the next page (this page's B)
};

template <>
struct ZFace<double> {
    double d[4];        // D0..D3 of the next page
    double w0, w1, w2;  // back-side weights of the next page
    double tA, tB;      // back-side t0 - t1 and t1 - t2 of the next page
    bool cross;         // v2 * v3 < 0 at the next page
};

// the fore side Q(d4, d3, d2, d1) of a page with B = d1 - 2 d2 + d3 and
// ic = 39/64 d2^2 + e'; leaves the next page's back-side values in f
__device__ __forceinline__ float zface_fore(ZFace<float>& f, float d1, float d2, float d3, float d4, float ic,
                                            float B, float e)
{
    const float ib = fmaf(0.609375f * d3, d3, e);
    const float u0 = fmaf(-1.375f, d3, d4), u1 = fmaf(-0.625f, d2, d3), u2 = fmaf(-1.375f, d2, d1);
    const float r0 = __builtin_amdgcn_rcpf(fmaf(u0, u0, ib));
    const float r1 = __builtin_amdgcn_rcpf(fmaf(u1, u1, ic));
    const float r2 = __builtin_amdgcn_rcpf(fmaf(u2, u2, ic));
    const float x0 = r0 * r0, x2 = r2 * r2;
    const float W = fmaf(6.f * r1, r1, fmaf(3.f, x2, x0));
    const float A = fmaf(-2.f, d3, d4) + d2;
    const float P = fmaf(4.f * x0, A, (6.f * x2) * B);
    f.d[0] = d1;
    f.d[1] = d2;
    f.d[2] = d3;
    f.d[3] = d4;
    f.ic = ib;
    f.x0 = x2;
    f.x2 = x0;
    f.A = B;
    return fmaf(P, __builtin_amdgcn_rcpf(__builtin_fmaxf(W, 1e-30f)), -B);
}

__device__ __forceinline__ void zface_start(ZFace<float>& f, const float p[7], JPf j)
{
    float d[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = fmaf(-2.f, p[k + 1], p[k]) + p[k + 2];
    const float ic = fmaf(0.609375f * d[1], d[1], j.e);
    (void)zface_fore(f, d[0], d[1], d[2], d[3], ic, fmaf(-2.f, d[1], d[0]) + d[2], j.e);
}

// weno5_pair on the page's z-line p[0..6] with the faces of the page before
__device__ __forceinline__ void weno5_pair_z(float& d_fore, float& d_back, const float p[7], ZFace<float>& f, JPf j)
{
    const float d1 = f.d[1], d2 = f.d[2], d3 = f.d[3];
    const float d4 = fmaf(-2.f, p[5], p[4]) + p[6];
    const float c = fmaf(8.f, p[4] - p[2], p[1] - p[5]);
    const float ic = f.ic;
    const float B = fmaf(-2.f, d2, d1) + d3;
    // back: a = d0, b = d1, c = d2, d = d3; 1/s0^2, 1/s2^2 and A from the page before
    float qb;
    {
        const float u1 = fmaf(-0.625f, d2, d1);
        const float r1 = __builtin_amdgcn_rcpf(fmaf(u1, u1, ic));
        const float W = fmaf(6.f * r1, r1, fmaf(3.f, f.x2, f.x0));
        const float P = fmaf(4.f * f.x0, f.A, (6.f * f.x2) * B);
        qb = fmaf(P, __builtin_amdgcn_rcpf(__builtin_fmaxf(W, 1e-30f)), -B);
    }
    const float qf = zface_fore(f, d1, d2, d3, d4, ic, B, j.e);
    d_back = (c - qb) * j.k;
    d_fore = (c + qf) * j.k;
}

// FP64: jp_q(d4, d3, d2, d1) (its exact operation sequence) with its weights
// and differences left in f as the next page's back side
__device__ __forceinline__ double zface_fore(ZFace<double>& f, double d1, double d2, double d3, double d4, double e)
{
    const double k = 3.0 / 13.0;
    const double t0 = d4 - d3, t1 = d3 - d2, t2 = d2 - d1;
    const double y0 = fma(-3.0, d3, d4), y1 = d3 + d2, y2 = fma(3.0, d2, -d1);
    const double s0 = fma(t0, t0, fma(k * y0, y0, e));
    const double s1 = fma(t1, t1, fma(k * y1, y1, e));
    const double s2 = fma(t2, t2, fma(k * y2, y2, e));
    const double a0 = s0 * s0, a1 = s1 * s1, a2 = s2 * s2;
    const double w0 = a1 * a2, m01 = a0 * a1, w1 = 6.0 * (a0 * a2), w2 = 3.0 * m01;
    const double W = (w0 + w1) + w2;
    const double tA = t0 - t1, tB = t1 - t2;
    f.d[0] = d1;
    f.d[1] = d2;
    f.d[2] = d3;
    f.d[3] = d4;
    // the next page's back side: s0, s1, s2 are this side's s2, s1, s0, so
    // w0 = a1 a0, w1 = 6 (a2 a0), w2 = 3 (a2 a1), t0 - t1 = tB, t1 - t2 = tA
    f.w0 = m01;
    f.w1 = w1;
    f.w2 = 3.0 * w0;
    f.tA = tB;
    f.tB = tA;
    return fma(4.0 * w0, tA, fma(2.0, w2, -W) * tB) / W;
}

__device__ __forceinline__ void zface_start(ZFace<double>& f, const double p[7], JPd j)
{
    double d[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = fma(-2.0, p[k + 1], p[k]) + p[k + 2];
    (void)zface_fore(f, d[0], d[1], d[2], d[3], j.e);
    f.cross = p[2] * p[3] < 0;
}

__device__ __forceinline__ void weno5_pair_z(double& d_fore, double& d_back, bool& cross, const double p[7],
                                             ZFace<double>& f, JPd j)
{
    const double d1 = f.d[1], d2 = f.d[2], d3 = f.d[3];
    const double d4 = fma(-2.0, p[5], p[4]) + p[6];
    const double c = fma(8.0, p[4] - p[2], p[1] - p[5]);
    const double W = (f.w0 + f.w1) + f.w2;
    const double qb = fma(4.0 * f.w0, f.tA, fma(2.0, f.w2, -W) * f.tB) / W;
    const bool up = p[3] * p[4] < 0;
    cross = f.cross | up;
    f.cross = up;
    const double qf = zface_fore(f, d1, d2, d3, d4, j.e);
    d_back = (c - qb) * j.k;
    d_fore = (c + qf) * j.k;
}

// ---------------------------------------------------------------------------
// Nonuniform three-stencil WENO at a node next to the interface:
// re_step.cu:79-243 (weno_nonuniform), reconstructing at face a[2] (fore) or
// a[3] (back) from first divided differences u[k] over [a[k], a[k+1]].
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ S weno_nonuniform_face(const S* a, const S* u, bool back)
{
#pragma clang fp contract(off)
    const S xm25 = a[0], xm15 = a[1], xm05 = a[2], x05 = a[3], x15 = a[4], x25 = a[5];
    const S um2 = u[0], um1 = u[1], u0 = u[2], u1 = u[3], u2 = u[4];
    S v0, v1, v2, c0, c1, c2;
    if (!back) {
        v0 = u1 + (u0 - u1) * (1.0 + (x05 - xm05) / (x15 - xm05) + (x05 - xm05) / (x25 - xm05))
                + (u2 - u1) * ((x05 - xm05) / (x25 - xm05)) * ((x15 - xm05) / (x25 - x05));
        v1 = u0 + (um1 - u0) * ((x05 - xm05) / (x15 - xm15)) * ((x15 - xm05) / (x05 - xm15))
                - (u1 - u0) * ((x05 - xm05) / (x15 - xm15)) * ((xm05 - xm15) / (x15 - xm05));
        v2 = um1 + (u0 - um1) * ((xm05 - xm15) / (x05 - xm25)) * ((xm05 - xm25) / (x05 - xm15))
                 - (um2 - um1) * ((xm05 - xm15) / (x05 - xm25)) * ((x05 - xm05) / (xm05 - xm25));
        c0 = ((xm05 - xm15) / (x25 - xm25)) * ((xm05 - xm25) / (x25 - xm15));
        c1 = ((xm05 - xm25) / (x25 - xm25)) * ((x25 - xm05) / (x25 - xm15))
             * (1.0 + (x25 - xm15) / (x15 - xm25));
        c2 = ((x15 - xm05) / (x25 - xm25)) * ((x25 - xm05) / (x15 - xm25));
    } else {
        v0 = u1 + (u0 - u1) * ((x15 - x05) / (x25 - xm05)) * ((x25 - x05) / (x15 - xm05))
                - (u2 - u1) * ((x15 - x05) / (x25 - xm05)) * ((x05 - xm05) / (x25 - x05));
        v1 = u0 + (u1 - u0) * ((x05 - xm05) / (x15 - xm15)) * ((x05 - xm15) / (x15 - xm05))
                - (um1 - u0) * ((x05 - xm05) / (x15 - xm15)) * ((x15 - x05) / (x05 - xm15));
        v2 = um1 + (um2 - um1) * ((x05 - xm05) / (xm05 - xm25)) * ((x05 - xm15) / (x05 - xm25))
                 + (u0 - um1) * (1.0 + (x05 - xm05) / (x05 - xm15) + (x05 - xm05) / (x05 - xm25));
        c0 = ((x05 - xm25) / (x25 - xm25)) * ((x05 - xm15) / (x25 - xm15));
        c1 = ((x05 - xm25) / (x25 - xm25)) * ((x25 - x05) / (x25 - xm15))
             * (1.0 + (x25 - xm15) / (x15 - xm25));
        c2 = ((x15 - x05) / (x25 - xm25)) * ((x25 - x05) / (x15 - xm25));
    }
    const S h2 = sq(x05 - xm05);
    const S s0 = 4.0 * sq((x05 - xm05) / (x25 - xm05)) *
        (sq((u2 - u1) / (x25 - x05)) * (10.0 * h2 + (x15 - xm05) * (x15 - x05))
         + (u2 - u1) * (u0 - u1) / ((x25 - x05) * (x15 - xm05)) *
               (20.0 * h2 + 2.0 * (x15 - xm05) * (x15 - x05) + (x25 - xm05) * (2.0 * x15 - x05 - xm05))
         + sq((u0 - u1) / (x15 - xm05)) * (10.0 * h2 + (x25 + x15 - 2.0 * xm05) * (x25 + x15 - x05 - xm05)));
    const S s1 = 4.0 * sq((x05 - xm05) / (x15 - xm15)) *
        (sq((um1 - u0) / (x05 - xm15)) * (10.0 * h2 + (x15 - xm05) * (x15 - x05))
         + (u1 - u0) * (um1 - u0) / ((x15 - xm05) * (x05 - xm15)) *
               (20.0 * h2 - (x15 - x05) * (xm05 - xm15) - (x15 - xm05) * (x05 - xm15))
         + sq((u1 - u0) / (x15 - xm05)) * (10.0 * h2 + (xm05 - xm15) * (x05 - xm15)));
    const S s2 = 4.0 * sq((x05 - xm05) / (x05 - xm25)) *
        (sq((um2 - um1) / (xm05 - xm25)) * (10.0 * h2 + (x05 - xm15) * (xm05 - xm15))
         + (u0 - um1) * (um2 - um1) / ((x05 - xm15) * (xm05 - xm25)) *
               (20.0 * h2 + 2.0 * (x05 - xm15) * (xm05 - xm15) + (x05 - xm25) * (x05 + xm05 - 2.0 * xm15))
         + sq((u0 - um1) / (x05 - xm15)) *
               (10.0 * h2 + (2.0 * x05 - xm25 - xm15) * (x05 + xm05 - xm15 - xm25)));
    const double eps = 1e-6;
    const S al0 = c0 / sq(s0 + eps), al1 = c1 / sq(s1 + eps), al2 = c2 / sq(s2 + eps);
    const S sum = al0 + al1 + al2;
    return v0 * (al0 / sum) + v1 * (al1 / sum) + v2 * (al2 / sum);
}

template <typename S>
struct PairT {
    S fore, back;
};
typedef PairT<double> Pair;

// select_stencil (re_step.cu:41-77; interface values h_fore/h_back as in
// 2_2_WENO5RK3_Extend/shared_utilities.cup:49-85) and the first divided
// differences (re_step.cu:83-88)
template <typename S>
struct NonuniformStencil {
    S x[7], d[6];
};

template <typename S>
__device__ __forceinline__ NonuniformStencil<S> nonuniform_stencil(S p0, S p1, S p2, S p3, S p4, S p5, S p6, S r1,
                                                                   S l1, double ds, S h_fore, S h_back)
{
#pragma clang fp contract(off)
    NonuniformStencil<S> st;
    S h[7];
    h[3] = p3;
    st.x[3] = S(0.0);
    if (val(r1) < ds) {
        st.x[4] = r1; st.x[5] = S(ds); st.x[6] = S(2 * ds);
        h[4] = h_fore; h[5] = p4; h[6] = p5;
    } else {
        st.x[4] = S(ds); st.x[5] = S(2 * ds); st.x[6] = S(3 * ds);
        h[4] = p4; h[5] = p5; h[6] = p6;
    }
    if (val(l1) < ds) {
        st.x[2] = -l1; st.x[1] = S(-ds); st.x[0] = S(-2 * ds);
        h[2] = h_back; h[1] = p2; h[0] = p1;
    } else {
        st.x[2] = S(-ds); st.x[1] = S(-2 * ds); st.x[0] = S(-3 * ds);
        h[2] = p2; h[1] = p1; h[0] = p0;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) st.d[k] = (h[k + 1] - h[k]) / (st.x[k + 1] - st.x[k]);
    return st;
}

// weno_nonuniform (re_step.cu:79-243), both faces.  The _inl forms are for
// kernels whose register budget is capped (the fused stage kernels, where
// the interface lanes run in blocks of their own); the others call the
// out-of-line forms: they run only at interface nodes, and scalar arguments
// and a returned struct keep the caller's stencil in registers.
template <typename S>
__device__ __forceinline__ PairT<S> weno_nonuniform_pair_inl(S p0, S p1, S p2, S p3, S p4, S p5, S p6, S r1, S l1,
                                                             double ds, S h_fore, S h_back)
{
    const NonuniformStencil<S> st = nonuniform_stencil(p0, p1, p2, p3, p4, p5, p6, r1, l1, ds, h_fore, h_back);
    PairT<S> out;
    out.fore = weno_nonuniform_face(&st.x[1], &st.d[1], false);
    out.back = weno_nonuniform_face(&st.x[0], &st.d[0], true);
    return out;
}
template <typename S>
__device__ __noinline__ PairT<S> weno_nonuniform_pair(S p0, S p1, S p2, S p3, S p4, S p5, S p6, S r1, S l1,
                                                      double ds, S h_fore, S h_back)
{
    return weno_nonuniform_pair_inl(p0, p1, p2, p3, p4, p5, p6, r1, l1, ds, h_fore, h_back);
}

// one face only: d_fore (back == false) or d_back
template <typename S>
__device__ __forceinline__ S weno_nonuniform_one_inl(S p0, S p1, S p2, S p3, S p4, S p5, S p6, S r1, S l1,
                                                     double ds, S h_fore, S h_back, bool back)
{
    const NonuniformStencil<S> st = nonuniform_stencil(p0, p1, p2, p3, p4, p5, p6, r1, l1, ds, h_fore, h_back);
    return back ? weno_nonuniform_face(&st.x[0], &st.d[0], true) : weno_nonuniform_face(&st.x[1], &st.d[1], false);
}
template <typename S>
__device__ __noinline__ S weno_nonuniform_one(S p0, S p1, S p2, S p3, S p4, S p5, S p6, S r1, S l1, double ds,
                                              S h_fore, S h_back, bool back)
{
    return weno_nonuniform_one_inl(p0, p1, p2, p3, p4, p5, p6, r1, l1, ds, h_fore, h_back, back);
}

template <typename S>
__device__ __forceinline__ PairT<S> weno_nonuniform_pair(const S v[7], S r1, S l1, double ds, S h_fore, S h_back)
{
    return weno_nonuniform_pair(v[0], v[1], v[2], v[3], v[4], v[5], v[6], r1, l1, ds, h_fore, h_back);
}

// ---------------------------------------------------------------------------
// Godunov Hamiltonian residual, re_step.cu:382-392
// ---------------------------------------------------------------------------
// Per axis the reference takes max(min(0,l)^2, max(0,r)^2) inside and
// max(max(0,l)^2, min(0,r)^2) outside.  The larger square is the square of
// the larger magnitude, so each is the square of one three-way extremum:
// min(0, l, -r) inside and max(0, l, -r) outside.  Squaring depends only on
// the magnitude, so the result is bit-identical to the reference's form with
// a third of the (half-rate) min/max operations.
template <typename S>
__device__ __forceinline__ S godunov_axis(bool inside, S l, S r)
{
    const S z(0.0);
    const S m = inside ? dmin(dmin(z, l), -r) : dmax(dmax(z, l), -r);
    return m * m;
}
template <typename S>
__device__ __forceinline__ S godunov_step(bool inside, S xR, S xL, S yF, S yB, S zU, S zD, S dt)
{
    const S g = godunov_axis(inside, xL, xR) + godunov_axis(inside, yB, yF) + godunov_axis(inside, zD, zU);
    const S s = (sqrt(g) - 1.0) * dt;
    return inside ? -s : s;
}

__device__ __forceinline__ float fsq(float a) { return a * a; }

// FP32 form of godunov_step (bulk points): the three-way extremum of
// godunov_axis as one v_min3 / v_max3 per axis
__device__ __forceinline__ float godunov_axis_f(bool inside, float l, float r)
{
    const float m = inside ? __builtin_fminf(__builtin_fminf(0.f, l), -r) : __builtin_fmaxf(__builtin_fmaxf(0.f, l), -r);
    return m * m;
}
__device__ __forceinline__ float godunov_step(bool inside, float xR, float xL, float yF, float yB, float zU,
                                              float zD, float dt)
{
    const float g = godunov_axis_f(inside, xL, xR) + godunov_axis_f(inside, yB, yF) + godunov_axis_f(inside, zD, zU);
    const float s = (__builtin_sqrtf(g) - 1.f) * dt;
    return inside ? -s : s;
}

// ---------------------------------------------------------------------------
// ENO2 one-sided derivatives with subcell node value 0,
// 1_0_ENORK2_Reinitialization/boundary_correction.cu:80-100
// ---------------------------------------------------------------------------
__device__ __forceinline__ void eno2_pair(double& sR, double& sL, const double p[7], double pr, double pl,
                                          double ds)
{
    const double v4 = p[1], v1 = p[2], v0 = p[3], v2 = p[4], v3 = p[5];
    const double p2 = v1 - 2.0 * v0 + v2;
    const double p2r = v0 - 2.0 * v2 + v3;
    double p2m = 0.5 * min_mod(p2, p2r) / (ds * ds);
    const double vr = (pr == ds) ? v2 : 0.0;
    sR = (vr - v0) / pr - pr * p2m;
    const double p2l = v0 - 2.0 * v1 + v4;
    p2m = 0.5 * min_mod(p2, p2l) / (ds * ds);
    const double vl = (pl == ds) ? v1 : 0.0;
    sL = (v0 - vl) / pl + pl * p2m;
}

// ---------------------------------------------------------------------------
// subcell crossing distances
// ---------------------------------------------------------------------------
// quadratic ENO: 1_3_WENO5RK3_Reinitialization/boundary_location.cu:6-20
__device__ __forceinline__ double eno_quadratic_distance(double v1, double v0, double v2, double v3, double ds)
{
#pragma clang fp contract(off)
    const double eps = 10e-10;
    const double p2l = v2 - 2.0 * v0 + v1;
    const double p2r = v3 - 2.0 * v2 + v0;
    const double p2 = min_mod(p2l, p2r);
    if (p2 > eps || p2 < -eps) {
        const double disc = sq(0.5 * p2 - v0 - v2) - 4. * v0 * v2;
        return ds * (0.5 + (v0 - v2 - ksign(v0 - v2) * sqrt(disc)) / p2);
    }
    return ds * v0 / (v0 - v2);
}

// Cubic crossing distance of the WENO-RK3 variants (cubic_distance,
// 1_1_WENORK3_Reinitialization/weno_rk3_reinitialization.cu:74-117): the
// cubic through (0,v0),(s,v1),(2s,v2),(3s,v3) with v1*v2 < 0, Newton from the
// linear guess (at most 50 steps, until |f/d| <= 1e-14), the root clamped to
// [s, 2s]; next to a kink the quadratic ENO distance (sp, :55-69).  The
// reference's pow(s,2) and pow(s,3) are s*s and s*s*s; f/d is formed once.
__device__ __forceinline__ double cubic_distance(double v0, double v1, double v2, double v3, double s)
{
#pragma clang fp contract(off)
    if (v0 * v1 < 0 || v2 * v3 < 0) return eno_quadratic_distance(v0, v1, v2, v3, s);
    const double c0 = v0;
    const double c1 = (3.0 * (v1 - v0) - 3.0 / 2.0 * (v2 - v0) + 1.0 / 3.0 * (v3 - v0)) / s;
    const double c2 = (-5.0 / 2.0 * (v1 - v0) + 2.0 * (v2 - v0) - 1.0 / 2.0 * (v3 - v0)) / (s * s);
    const double c3 = (1.0 / 2.0 * (v1 - v0) - 1.0 / 2.0 * (v2 - v0) + 1.0 / 6.0 * (v3 - v0)) / (s * s * s);
    double xc = s + s * v1 / (v1 - v2);
    double diff = 1;
    for (int iter = 0; diff > 1e-14 && iter < 50; ++iter) {
        const double f = c0 + c1 * xc + c2 * xc * xc + c3 * xc * xc * xc;
        const double d = c1 + 2 * c2 * xc + 3 * c3 * xc * xc;
        const double q = f / d;
        diff = fabs(q);
        xc = xc - q;
    }
    xc = dmax(xc, s);
    xc = dmin(xc, 2 * s);
    return xc - s;
}

// Lagrange -> monomial coefficients of the quintic through (x_i, y_i)
// (Poly<5>::setInterpolation, shared_utilities.hpp:57-103).  The monomial form
// is ill-conditioned (~1e-11 relative near the data), so this and the
// evaluation below keep one operation order without contraction; interface
// cells are few, the cost does not matter.
__device__ __forceinline__ void poly5_interp(const double* x, const double* y, double coef[6])
{
#pragma clang fp contract(off)
    double c[6][6];
    double kk[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double k = y[i];
        c[i][0] = 1.0;
#pragma unroll
        for (int m = 1; m < 6; ++m) c[i][m] = 0.0;
        int deg = 0;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            if (j == i) continue;
            k /= x[i] - x[j];
            // multiply c[i] (degree deg) by (x - x[j])
#pragma unroll
            for (int m = 5; m >= 1; --m)
                if (m <= deg + 1) c[i][m] = c[i][m] - x[j] * c[i][m - 1];
            ++deg;
        }
        kk[i] = k;
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) acc += kk[j] * c[j][5 - m];
        coef[m] = acc;
    }
}

// sum_i c_i x^i in index order (Poly<n>::operator(), shared_utilities.hpp:43-49)
__device__ __forceinline__ double poly5_eval(const double coef[6], double x)
{
#pragma clang fp contract(off)
    double v = 0.0, p = 1.0;
#pragma unroll
    for (int i = 0; i <= 5; ++i) {
        v += coef[i] * p;
        p *= x;
    }
    return v;
}

// Brent root of the quintic in [x1,x2] with the reference's control flow
// (zeroin, shared_utilities.hpp:109-173; abs() of doubles is fabs())
__device__ __forceinline__ double zeroin_poly5(const double* coef, double x1, double x2, double toler)
{
#pragma clang fp contract(off)
    const int ITMAX = 100;
    double d = 0, e = 0, p, q, r, s, xm;
    double a = x1, b = x2, c = x2;
    double fa = poly5_eval(coef, a), fb = poly5_eval(coef, b), fc = fb;
    for (int it = 0; it < ITMAX; ++it) {
        if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) {
            c = a; fc = fa;
            d = b - a; e = d;
        }
        if (fabs(fc) < fabs(fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        xm = 0.5 * (c - b);
        if (fabs(xm) <= toler || fb == 0.0) return b;
        if (fabs(e) < toler || fabs(fa) <= fabs(fb)) {
            d = xm; e = xm;
        } else {
            s = fb / fa;
            if (a == c) {
                p = 2.0 * xm * s;
                q = 1.0 - s;
            } else {
                q = fa / fc;
                r = fb / fc;
                p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0) q = -q;
            p = fabs(p);
            const double m1 = 3.0 * xm * q - fabs(toler * q);
            const double m2 = fabs(e * q);
            if (2.0 * p < dmin(m1, m2)) {
                e = d; d = p / q;
            } else {
                d = xm; e = d;
            }
        }
        a = b; fa = fb;
        if (fabs(d) > toler) b = b + d;
        else b += ksign(xm) * toler;
        fb = poly5_eval(coef, b);
    }
    return b;
}

// six_distance: boundary_location.cu:53-80
__device__ __noinline__ double six_distance(double v0, double v1, double v2, double v3, double v4, double v5,
                                            double s)
{
#pragma clang fp contract(off)
    if (v1 * v2 < 0 || v3 * v4 < 0) return eno_quadratic_distance(v1, v2, v3, v4, s);
    const double xx[6] = {0.0, s, 2 * s, 3 * s, 4 * s, 5 * s};
    const double yy[6] = {v0, v1, v2, v3, v4, v5};
    double coef[6];
    poly5_interp(xx, yy, coef);
    double xc = zeroin_poly5(coef, 2 * s, 3 * s, 3e-16);
    xc = dmax(xc, 2 * s);
    xc = dmin(xc, 3 * s);
    return xc - 2 * s;
}

// sixth_interp: 2_2_WENO5RK3_Extend/boundary_interpolant.cu:77-99
__device__ __noinline__ Pair sixth_interp(double dis_f, double dis_b, double s, double v0, double v1, double v2,
                                          double v3, double v4, double v5, double v6)
{
    Pair out{0.0, 0.0};
    if (dis_f < s) {
        const double xf[6] = {s, 2 * s, 3 * s, 4 * s, 5 * s, 6 * s};
        const double yf[6] = {v1, v2, v3, v4, v5, v6};
        double coef[6];
        poly5_interp(xf, yf, coef);
        out.fore = poly5_eval(coef, 3 * s + dis_f);
    }
    if (dis_b < s) {
        const double xb[6] = {0.0, s, 2 * s, 3 * s, 4 * s, 5 * s};
        const double yb[6] = {v0, v1, v2, v3, v4, v5};
        double coef[6];
        poly5_interp(xb, yb, coef);
        out.back = poly5_eval(coef, 3 * s - dis_b);
    }
    return out;
}

// ---------------------------------------------------------------------------
// WENO-RK3 and ENO-RK2 variants of the reinitialization and extension
// (1_1_WENORK3_Reinitialization, 2_0_ENORK2_Extend, 2_1_WENORK3_Extend)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double min_abs(double x, double y) { return (fabs(x) < fabs(y)) ? x : y; }

// select_stencil (weno_rk3_reinitialization.cu:149-184, with the interface
// values h_fore/h_back of wenork3_extend.cu:88-123): nodes x[0..6] = x3m..x3
// (x[3] = 0 the point) and values h[0..6] from p[0..6], the crossing taking
// the place of the first node on a side whose distance is below ds
__device__ __forceinline__ void cubic_stencil(double x[7], double h[7], const double p[7], double r1, double l1,
                                              double ds, double h_fore, double h_back)
{
    x[3] = 0.0;
    h[3] = p[3];
    if (r1 < ds) {
        x[4] = r1; x[5] = ds; x[6] = 2 * ds;
        h[4] = h_fore; h[5] = p[4]; h[6] = p[5];
    } else {
        x[4] = ds; x[5] = 2 * ds; x[6] = 3 * ds;
        h[4] = p[4]; h[5] = p[5]; h[6] = p[6];
    }
    if (l1 < ds) {
        x[2] = -l1; x[1] = -ds; x[0] = -2 * ds;
        h[2] = h_back; h[1] = p[2]; h[0] = p[1];
    } else {
        x[2] = -ds; x[1] = -2 * ds; x[0] = -3 * ds;
        h[2] = p[2]; h[1] = p[1]; h[0] = p[0];
    }
}

// ENO_cubic_derivative (weno_rk3_reinitialization.cu:190-222, the same text in
// wenork3_extend.cu:129-164): cubic ENO one-sided derivatives at x[3] on the
// nonuniform stencil, from Newton divided differences
__device__ __forceinline__ void eno_cubic_pair(double& d_fore, double& d_back, const double x[7], const double h[7])
{
#pragma clang fp contract(off)
    const double d1_25 = (h[6] - h[5]) / (x[6] - x[5]);
    const double d1_15 = (h[5] - h[4]) / (x[5] - x[4]);
    const double d1_05 = (h[4] - h[3]) / (x[4] - x[3]);
    const double d1_m05 = (h[3] - h[2]) / (x[3] - x[2]);
    const double d1_m15 = (h[2] - h[1]) / (x[2] - x[1]);
    const double d1_m25 = (h[1] - h[0]) / (x[1] - x[0]);
    const double d2_2 = (d1_25 - d1_15) / (x[6] - x[4]);
    const double d2_1 = (d1_15 - d1_05) / (x[5] - x[3]);
    const double d2_0 = (d1_05 - d1_m05) / (x[4] - x[2]);
    const double d2_m1 = (d1_m05 - d1_m15) / (x[3] - x[1]);
    const double d2_m2 = (d1_m15 - d1_m25) / (x[2] - x[0]);
    const double d3_15 = (d2_2 - d2_1) / (x[6] - x[3]);
    const double d3_05 = (d2_1 - d2_0) / (x[5] - x[2]);
    const double d3_m05 = (d2_0 - d2_m1) / (x[4] - x[1]);
    const double d3_m15 = (d2_m1 - d2_m2) / (x[3] - x[0]);
    const double a1 = (x[3] - x[2]) * (x[3] - x[1]) * min_abs(d3_m05, d3_m15);
    const double a2 = (x[3] - x[2]) * (x[3] - x[4]) * min_abs(d3_m05, d3_05);
    const double a = (fabs(d2_m1) < fabs(d2_0)) ? a1 : a2;
    const double b1 = (x[3] - x[2]) * (x[3] - x[4]) * min_abs(d3_m05, d3_05);
    const double b2 = (x[3] - x[4]) * (x[3] - x[5]) * min_abs(d3_05, d3_15);
    const double b = (fabs(d2_0) < fabs(d2_1)) ? b1 : b2;
    d_back = d1_m05 + min_mod(d2_m1, d2_0) * (x[3] - x[2]) + a;
    d_fore = d1_05 + min_mod(d2_0, d2_1) * (x[3] - x[4]) + b;
}

// weno_derivative_boundary of the WENO-RK3 variants: cubic ENO on the
// interface stencil where `cross`, else the uniform WENO5 pair (the same
// weno_onesided_derivative as re_step.cu:11-37, in the Jiang-Peng form).
// The reinitialization tests a sign change of p (weno_rk3_reinitialization.cu:228-247,
// interface values 0), the extension a subcell distance (wenork3_extend.cu:167-190).
__device__ __forceinline__ void wenork3_pair(double& fore, double& back, const double p[7], double r1, double l1,
                                             double ds, JPd jp, bool cross, double h_fore, double h_back)
{
    if (cross) {
        double x[7], h[7];
        cubic_stencil(x, h, p, r1, l1, ds, h_fore, h_back);
        eno_cubic_pair(fore, back, x, h);
    } else {
        weno5_pair(fore, back, p, jp);
    }
}

// eno_derivative_field (enork2_extend.cu:56-78): ENO2 one-sided derivatives
// with the interface values v_fore/v_back where a distance differs from ds
__device__ __forceinline__ void eno2_field_pair(double& sR, double& sL, const double p[7], double dis_f, double dis_b,
                                                double ds, double v_fore, double v_back)
{
#pragma clang fp contract(off)
    const double v4 = p[1], v1 = p[2], v0 = p[3], v2 = p[4], v3 = p[5];
    const double p2 = v1 - 2.0 * v0 + v2;
    const double p2r = v0 - 2.0 * v2 + v3;
    double p2m = 0.5 * min_mod(p2, p2r) / (ds * ds);
    const double v_f = (dis_f == ds) ? v2 : v_fore;
    sR = (v_f - v0) / dis_f - dis_f * p2m;
    const double p2l = v0 - 2.0 * v1 + v4;
    p2m = 0.5 * min_mod(p2, p2l) / (ds * ds);
    const double v_b = (dis_b == ds) ? v1 : v_back;
    sL = (v0 - v_b) / dis_b + dis_b * p2m;
}

// ENO2 interface values (cubic_eno_interp, enork2_extend.cu:144-170): the
// quadratic ENO interpolant through the two nodes around a crossing, at the
// crossing, for a side whose distance differs from ds (0 otherwise)
__device__ __forceinline__ void eno2_interface_values(double& c_fore, double& c_back, double dis_f, double dis_b,
                                                      double ds, const double p[7])
{
#pragma clang fp contract(off)
    const double v4 = p[1], v1 = p[2], v0 = p[3], v2 = p[4], v3 = p[5];
    c_fore = 0;
    c_back = 0;
    if (dis_f != ds) {
        const double p2 = v2 - 2.0 * v0 + v1;
        const double p2r = v3 - 2.0 * v2 + v0;
        const double c2 = 0.5 * min_mod(p2, p2r) / (ds * ds);
        const double c1 = (v2 - v0) / ds;
        const double c0 = (v2 + v0) / 2 - c2 * (ds * ds) / 4;
        const double x = dis_f - ds / 2;
        c_fore = c2 * x * x + c1 * x + c0;
    }
    if (dis_b != ds) {
        const double p2 = v2 - 2.0 * v0 + v1;
        const double p2l = v0 - 2.0 * v1 + v4;
        const double c2 = 0.5 * min_mod(p2, p2l) / (ds * ds);
        const double c1 = (v0 - v1) / ds;
        const double c0 = (v1 + v0) / 2 - c2 * (ds * ds) / 4;
        const double x = ds / 2 - dis_b;
        c_back = c2 * x * x + c1 * x + c0;
    }
}

// cubic interface values (cubic_interp, wenork3_extend.cu:310-329): the cubic
// through the four nodes around a crossing (cubic_interp_coefficient, :297-306),
// at the crossing, for a side whose distance is below ds (0 otherwise).  The
// reference's pow(xc,2), pow(xc,3) are xc*xc and xc*xc*xc.
__device__ __forceinline__ void cubic_coefficients(double c[4], double v0, double v1, double v2, double v3, double s)
{
#pragma clang fp contract(off)
    c[0] = v0;
    c[1] = (3.0 * (v1 - v0) - 3.0 / 2.0 * (v2 - v0) + 1.0 / 3.0 * (v3 - v0)) / s;
    c[2] = (-5.0 / 2.0 * (v1 - v0) + 2.0 * (v2 - v0) - 1.0 / 2.0 * (v3 - v0)) / (s * s);
    c[3] = (1.0 / 2.0 * (v1 - v0) - 1.0 / 2.0 * (v2 - v0) + 1.0 / 6.0 * (v3 - v0)) / (s * s * s);
}
__device__ __forceinline__ void cubic_interface_values(double& c_fore, double& c_back, double dis_f, double dis_b,
                                                       double ds, const double p[7])
{
#pragma clang fp contract(off)
    c_fore = 0.;
    c_back = 0.;
    double c[4];
    if (dis_f < ds) {
        cubic_coefficients(c, p[2], p[3], p[4], p[5], ds);
        const double xc = ds + dis_f;
        c_fore = c[0] + c[1] * xc + c[2] * (xc * xc) + c[3] * (xc * xc * xc);
    }
    if (dis_b < ds) {
        cubic_coefficients(c, p[1], p[2], p[3], p[4], ds);
        const double xc = 2 * ds - dis_b;
        c_back = c[0] + c[1] * xc + c[2] * (xc * xc) + c[3] * (xc * xc * xc);
    }
}

}  // namespace vls
