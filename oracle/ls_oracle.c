/*
 * ls_oracle.c -- CPU oracle for the level-set hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product is
 * compared against.  It is loaded by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py, and by nothing else; the product library
 * never links, loads or calls it.
 *
 * Every function restates, in plain C, the arithmetic of one reference kernel
 * or MATLAB wrapper of tiankuizhang/2018_7_Vesicle and cites the file:line it
 * follows.  Paths are relative to src/7_0_paper/2_0_Vesicle/ (the canonical
 * snapshot) unless they start with src/.
 *
 * Conventions kept from the reference:
 *   - fields are double[rows*cols*pges], column-major [rows,cols,pges] with
 *     ind = pge*rows*cols + col*rows + row (CUDA_Code/1_3_WENO5RK3_Reinitialization/
 *     shared_utilities.hpp:21-31); rows = y, cols = x, pges = z;
 *   - every stencil neighbour index is clamped to the box;
 *   - kernel sign(x) is +1 for x>0 and -1 otherwise (shared_utilities.hpp:17-18),
 *     MATLAB sign() is 0 at 0 (used only by the extension wrappers).
 * Deliberate arithmetic differences (documented, tolerance-level only):
 *   - pow(x,2) is written x*x and Poly<n>'s pow(x,i) is repeated multiplication;
 *   - build with -ffp-contract=off so no fused multiply-adds are formed.
 *
 * Parity pins (no recorded outputs are needed): tests/test_oracle_math.py and
 * tests/test_oracle_kats.py (closed forms, known answers on kinked, linear,
 * polynomial and planar fields, orders of accuracy, invariants) and the
 * reference's own recorded runs (tests/test_oracle_relax.py).
 * tests/test_oracle_mutants.py plants a mistake in every function of this
 * file and shows that one of those checks fails.  The reference's CUDA
 * kernels are not built: they need nvcc, which this image lacks.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t idx_t;

/* ------------------------------------------------------------------------- */
/* small helpers: shared_utilities.hpp:4-18                                   */
/* ------------------------------------------------------------------------- */
static inline double dmax2(double x, double y) { return (x < y) ? y : x; }
static inline double dmin2(double x, double y) { return (x < y) ? x : y; }
static inline double sq(double x) { return x * x; }
static inline double min_mod(double x, double y)
{
    return (x * y < 0) ? 0.0 : (fabs(x) < fabs(y) ? x : y);
}
static inline double ksign(double x) { return (x > 0) ? 1.0 : -1.0; }
static inline double msign(double x) { return (x > 0) ? 1.0 : ((x < 0) ? -1.0 : 0.0); }
static inline int iclamp(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

/* clamped linear index, shared_utilities.hpp:21-31 */
static inline idx_t at(int r, int c, int p, int rows, int cols, int pges)
{
    r = iclamp(r, rows - 1);
    c = iclamp(c, cols - 1);
    p = iclamp(p, pges - 1);
    return (idx_t)p * rows * cols + (idx_t)c * rows + r;
}

/* The 7 values along one axis centred on (r,c,p): out[0..6] = offsets -3..+3.
 * axis 0 = x (cols), 1 = y (rows), 2 = z (pges). */
static inline void line7(const double *f, int r, int c, int p, int axis,
                         int rows, int cols, int pges, double out[7])
{
    for (int s = -3; s <= 3; ++s) {
        int rr = r, cc = c, pp = p;
        if (axis == 0) cc += s; else if (axis == 1) rr += s; else pp += s;
        out[s + 3] = f[at(rr, cc, pp, rows, cols, pges)];
    }
}

/* ------------------------------------------------------------------------- */
/* Poly<5> Lagrange -> monomial coefficients, shared_utilities.hpp:37-103     */
/* ------------------------------------------------------------------------- */
#define PN 5
static void poly5_interp(const double *x, const double *y, double coef[PN + 1])
{
    double kk[PN + 1], zeros[PN + 1][PN];
    for (int i = 0; i <= PN; ++i) {
        kk[i] = y[i];
        int m = 0;
        for (int j = 0; j <= PN; ++j) {
            if (j == i) continue;
            kk[i] /= x[i] - x[j];
            zeros[i][m++] = x[j];
        }
    }
    /* c[i][k]: coefficient of x^(deg-k) of prod_j (x - zeros[i][j]) */
    double c[PN + 1][PN + 1];
    for (int i = 0; i <= PN; ++i) {
        c[i][0] = 1.0;
        for (int k = 1; k <= PN; ++k) c[i][k] = 0.0;
        for (int j = 0; j < PN; ++j) {
            double t[PN + 1];
            for (int k = 1; k <= j + 1; ++k) t[k] = c[i][k] - zeros[i][j] * c[i][k - 1];
            for (int k = 1; k <= j + 1; ++k) c[i][k] = t[k];
        }
    }
    for (int m = 0; m <= PN; ++m) {
        double acc = 0.0;
        for (int j = 0; j <= PN; ++j) acc += kk[j] * c[j][PN - m];
        coef[m] = acc;
    }
}

/* Poly<n>::operator(), shared_utilities.hpp:43-49: sum_i c_i x^i in index order */
static double poly5_eval(const double coef[PN + 1], double x)
{
    double val = 0.0, p = 1.0;
    for (int i = 0; i <= PN; ++i) {
        val += coef[i] * p;
        p *= x;
    }
    return val;
}

/* Brent's zeroin as written in shared_utilities.hpp:109-173 (same branch
 * structure and tie-breaking; abs() on doubles is fabs()). */
static double zeroin_poly5(const double coef[PN + 1], double x1, double x2, double toler)
{
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
            double m1 = 3.0 * xm * q - fabs(toler * q);
            double m2 = fabs(e * q);
            if (2.0 * p < dmin2(m1, m2)) {
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

/* ------------------------------------------------------------------------- */
/* subcell interface distances                                                */
/* ------------------------------------------------------------------------- */

/* quadratic-ENO crossing distance from v0 toward v2 (v1 = node behind v0,
 * v3 = node beyond v2): CUDA_Code/1_3_WENO5RK3_Reinitialization/boundary_location.cu:6-20,
 * identical to CUDA_Code/1_0_ENORK2_Reinitialization/boundary_correction.cu:60-75 */
double or_eno_quadratic_distance(double v1, double v0, double v2, double v3, double ds)
{
    const double eps = 10e-10;
    double p2l = v2 - 2.0 * v0 + v1;
    double p2r = v3 - 2.0 * v2 + v0;
    double p2 = min_mod(p2l, p2r);
    if (p2 > eps || p2 < -eps) {
        double disc = sq(0.5 * p2 - v0 - v2) - 4. * v0 * v2;
        return ds * (0.5 + (v0 - v2 - ksign(v0 - v2) * sqrt(disc)) / p2);
    }
    return ds * v0 / (v0 - v2);
}

/* degree-5 Lagrange interpolant through 6 equispaced values + Brent root in
 * [2s,3s]; quadratic ENO near a kink: boundary_location.cu:53-80 */
double or_six_distance(double v0, double v1, double v2, double v3, double v4, double v5, double s)
{
    if (v1 * v2 < 0 || v3 * v4 < 0) return or_eno_quadratic_distance(v1, v2, v3, v4, s);
    double xx[6] = {0.0, s, 2 * s, 3 * s, 4 * s, 5 * s};
    double yy[6] = {v0, v1, v2, v3, v4, v5};
    double coef[6];
    poly5_interp(xx, yy, coef);
    double xc = zeroin_poly5(coef, 2 * s, 3 * s, 3e-16);
    xc = dmax2(xc, 2 * s);
    xc = dmin2(xc, 3 * s);
    return xc - 2 * s;
}

/* cubic crossing distance of the WENO-RK3 variants:
 * CUDA_Code/1_1_WENORK3_Reinitialization/weno_rk3_reinitialization.cu:74-117.  The cubic
 * through (0,v0),(s,v1),(2s,v2),(3s,v3), v1*v2<0, Newton from the linear
 * guess until |f/d| <= 1e-14 (at most 50 steps), clamped to [s,2s]; quadratic
 * ENO (sp, :55-69, argument order (v0,v1,v2,v3)) next to a kink.  pow(s,2),
 * pow(s,3) are s*s and s*s*s; f/d is formed once per step. */
double or_cubic_distance(double v0, double v1, double v2, double v3, double s)
{
    if (v0 * v1 < 0 || v2 * v3 < 0) return or_eno_quadratic_distance(v0, v1, v2, v3, s);
    double c0 = v0;
    double c1 = (3.0 * (v1 - v0) - 3.0 / 2.0 * (v2 - v0) + 1.0 / 3.0 * (v3 - v0)) / s;
    double c2 = (-5.0 / 2.0 * (v1 - v0) + 2.0 * (v2 - v0) - 1.0 / 2.0 * (v3 - v0)) / (s * s);
    double c3 = (1.0 / 2.0 * (v1 - v0) - 1.0 / 2.0 * (v2 - v0) + 1.0 / 6.0 * (v3 - v0)) / (s * s * s);
    double xc = s + s * v1 / (v1 - v2);
    double diff = 1;
    for (int iter = 0; diff > 1e-14 && iter < 50; ++iter) {
        double f = c0 + c1 * xc + c2 * xc * xc + c3 * xc * xc * xc;
        double d = c1 + 2 * c2 * xc + 3 * c3 * xc * xc;
        double q = f / d;
        diff = fabs(q);
        xc = xc - q;
    }
    xc = dmax2(xc, s);
    xc = dmin2(xc, 2 * s);
    return xc - s;
}

/* Kernel boundary_location (boundary_location.cu:84-132), when order==2
 * boundary_correction (CUDA_Code/1_0_ENORK2_Reinitialization/boundary_correction.cu:103-145),
 * when order==3 the cubic boundary_correction
 * (CUDA_Code/1_1_WENORK3_Reinitialization/weno_rk3_reinitialization.cu:251-292).
 * The six distance arrays are in/out: the caller pre-fills them with dx/dy/dz
 * (+SD/@SDF3/setDistance.m:5-10).  Each crossing writes the right-hand value
 * at ind and the left-hand value of the neighbour, exactly as the reference
 * (one writer per element, so the loop order does not matter). */
static void distances_impl(int order, double *xpr, double *xpl, double *ypf, double *ypb,
                           double *zpu, double *zpd, const double *lsf,
                           int rows, int cols, int pges, double dx, double dy, double dz)
{
    double *right_out[3] = {xpr, ypf, zpu};
    double *left_out[3] = {xpl, ypb, zpd};
    const double spacing[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                double f0 = lsf[ind];
                for (int axis = 0; axis < 3; ++axis) {
                    double v[7];
                    line7(lsf, r, c, p, axis, rows, cols, pges, v);
                    double f2 = v[4];
                    if (!(f0 * f2 < 0)) continue;
                    int rr = r, cc = c, pp = p;
                    if (axis == 0) cc += 1; else if (axis == 1) rr += 1; else pp += 1;
                    idx_t nb = at(rr, cc, pp, rows, cols, pges);
                    double ds = spacing[axis];
                    double d = (order == 2)   ? or_eno_quadratic_distance(v[2], f0, f2, v[5], ds)
                               : (order == 3) ? or_cubic_distance(v[2], f0, f2, v[5], ds)
                                              : or_six_distance(v[1], v[2], f0, f2, v[5], v[6], ds);
                    right_out[axis][ind] = d;
                    left_out[axis][nb] = ds - d;
                }
            }
}

void or_boundary_location(double *xpr, double *xpl, double *ypf, double *ypb, double *zpu,
                          double *zpd, const double *lsf, int num_ele, int rows, int cols,
                          int pges, double dx, double dy, double dz)
{
    (void)num_ele;
    distances_impl(6, xpr, xpl, ypf, ypb, zpu, zpd, lsf, rows, cols, pges, dx, dy, dz);
}

void or_boundary_correction(double *xpr, double *xpl, double *ypf, double *ypb, double *zpu,
                            double *zpd, const double *lsf, int num_ele, int rows, int cols,
                            int pges, double dx, double dy, double dz)
{
    (void)num_ele;
    distances_impl(2, xpr, xpl, ypf, ypb, zpu, zpd, lsf, rows, cols, pges, dx, dy, dz);
}

void or_cubic_boundary_correction(double *xpr, double *xpl, double *ypf, double *ypb, double *zpu,
                                  double *zpd, const double *lsf, int num_ele, int rows, int cols,
                                  int pges, double dx, double dy, double dz)
{
    (void)num_ele;
    distances_impl(3, xpr, xpl, ypf, ypb, zpu, zpd, lsf, rows, cols, pges, dx, dy, dz);
}

/* ------------------------------------------------------------------------- */
/* WENO5 derivatives                                                          */
/* ------------------------------------------------------------------------- */

/* The stencil-weight WENO5 one-sided derivative from 5 divided differences
 * (a check of or_weno_pair below, which the algorithms use),
 * CUDA_Code/1_3_WENO5RK3_Reinitialization/re_step.cu:11-37 (same text in
 * 2_2_WENO5RK3_Extend/shared_utilities.cup:19-45 and
 * 5_0_Nonoscillating_Reconstruction/nonoscillating_interpolant.cu:31-57).
 * The reference's alpha_k = c_k/(S_k+eps)^2, omega_k = alpha_k/sum(alpha),
 * d = sum omega_k phi_k is evaluated here as the algebraically identical
 * d = sum_k c_k phi_k prod_{j!=k} a_j / sum_k c_k prod_{j!=k} a_j with
 * a_k = (S_k+eps)^2.  The reference's own order is or_weno_onesided_reference
 * below; both are compared with or_weno_pair in tests/test_oracle_math.py. */
double or_weno_onesided(double v1, double v2, double v3, double v4, double v5)
{
    const double c13 = 1. / 3., c76 = 7. / 6., c116 = 11. / 6., c16 = 1. / 6., c56 = 5. / 6., c1312 = 13. / 12.;
    double phi1 = fma(c116, v3, fma(-c76, v2, c13 * v1));
    double phi2 = fma(c13, v4, fma(c56, v3, -c16 * v2));
    double phi3 = fma(-c16, v5, fma(c56, v4, c13 * v3));
    double d1 = fma(-2.0, v2, v1) + v3, e1 = fma(3.0, v3, fma(-4.0, v2, v1));
    double d2 = fma(-2.0, v3, v2) + v4, e2 = v2 - v4;
    double d3 = fma(-2.0, v4, v3) + v5, e3 = fma(3.0, v3, fma(-4.0, v4, v5));
    double S1 = fma(c1312 * d1, d1, 0.25 * e1 * e1);
    double S2 = fma(c1312 * d2, d2, 0.25 * e2 * e2);
    double S3 = fma(c1312 * d3, d3, 0.25 * e3 * e3);
    const double eps = 1e-6;
    double s1 = S1 + eps, s2 = S2 + eps, s3 = S3 + eps;
    double a1 = s1 * s1, a2 = s2 * s2, a3 = s3 * s3;
    double w1 = 0.1 * (a2 * a3), w2 = 0.6 * (a1 * a3), w3 = 0.3 * (a1 * a2);
    return fma(w1, phi1, fma(w2, phi2, w3 * phi3)) / (w1 + w2 + w3);
}

/* the reference's evaluation order, re_step.cu:11-37 (pow(x,2) written x*x) */
double or_weno_onesided_reference(double v1, double v2, double v3, double v4, double v5)
{
    double phi1 = 1. / 3. * v1 - 7. / 6. * v2 + 11. / 6. * v3;
    double phi2 = -1. / 6. * v2 + 5. / 6. * v3 + 1. / 3. * v4;
    double phi3 = 1. / 3. * v3 + 5. / 6. * v4 - 1. / 6. * v5;
    double S1 = 13. / 12. * sq(v1 - 2 * v2 + v3) + 1. / 4. * sq(v1 - 4 * v2 + 3 * v3);
    double S2 = 13. / 12. * sq(v2 - 2 * v3 + v4) + 1. / 4. * sq(v2 - v4);
    double S3 = 13. / 12. * sq(v3 - 2 * v4 + v5) + 1. / 4. * sq(3 * v3 - 4 * v4 + v5);
    const double eps = 1e-6;
    double a1 = 0.1 / sq(S1 + eps);
    double a2 = 0.6 / sq(S2 + eps);
    double a3 = 0.3 / sq(S3 + eps);
    double sum = a1 + a2 + a3;
    double w1 = a1 / sum, w2 = a2 / sum, w3 = a3 / sum;
    return w1 * phi1 + w2 * phi2 + w3 * phi3;
}

/* Nonuniform three-stencil WENO reconstruction of the derivative at a node
 * from first divided differences on cells bounded by the faces a[0..5]
 * (u[k] = divided difference over [a[k],a[k+1]]).  back==0 reconstructs at
 * face a[2] (d_fore), back==1 at face a[3] (d_back).
 * Formulas of re_step.cu:96-168 (fore) and :170-242 (back); the smoothness
 * indicators are the same expressions in both halves. */
static double weno_nonuniform_face(const double a[6], const double u[5], int back)
{
    const double xm25 = a[0], xm15 = a[1], xm05 = a[2], x05 = a[3], x15 = a[4], x25 = a[5];
    const double um2 = u[0], um1 = u[1], u0 = u[2], u1 = u[3], u2 = u[4];
    double v0, v1, v2, c0, c1, c2;
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
    const double h2 = sq(x05 - xm05);
    double s0 = 4.0 * sq((x05 - xm05) / (x25 - xm05)) *
                (sq((u2 - u1) / (x25 - x05)) * (10.0 * h2 + (x15 - xm05) * (x15 - x05))
                 + (u2 - u1) * (u0 - u1) / ((x25 - x05) * (x15 - xm05)) *
                       (20.0 * h2 + 2.0 * (x15 - xm05) * (x15 - x05)
                        + (x25 - xm05) * (2.0 * x15 - x05 - xm05))
                 + sq((u0 - u1) / (x15 - xm05)) *
                       (10.0 * h2 + (x25 + x15 - 2.0 * xm05) * (x25 + x15 - x05 - xm05)));
    double s1 = 4.0 * sq((x05 - xm05) / (x15 - xm15)) *
                (sq((um1 - u0) / (x05 - xm15)) * (10.0 * h2 + (x15 - xm05) * (x15 - x05))
                 + (u1 - u0) * (um1 - u0) / ((x15 - xm05) * (x05 - xm15)) *
                       (20.0 * h2 - (x15 - x05) * (xm05 - xm15) - (x15 - xm05) * (x05 - xm15))
                 + sq((u1 - u0) / (x15 - xm05)) * (10.0 * h2 + (xm05 - xm15) * (x05 - xm15)));
    double s2 = 4.0 * sq((x05 - xm05) / (x05 - xm25)) *
                (sq((um2 - um1) / (xm05 - xm25)) * (10.0 * h2 + (x05 - xm15) * (xm05 - xm15))
                 + (u0 - um1) * (um2 - um1) / ((x05 - xm15) * (xm05 - xm25)) *
                       (20.0 * h2 + 2.0 * (x05 - xm15) * (xm05 - xm15)
                        + (x05 - xm25) * (x05 + xm05 - 2.0 * xm15))
                 + sq((u0 - um1) / (x05 - xm15)) *
                       (10.0 * h2 + (2.0 * x05 - xm25 - xm15) * (x05 + xm05 - xm15 - xm25)));
    const double eps = 1e-6;
    double a0 = c0 / sq(s0 + eps), a1 = c1 / sq(s1 + eps), a2 = c2 / sq(s2 + eps);
    double sum = a0 + a1 + a2;
    double w0 = a0 / sum, w1 = a1 / sum, w2 = a2 / sum;
    return v0 * w0 + v1 * w1 + v2 * w2;
}

/* select_stencil + weno_nonuniform: re_step.cu:41-77,79-243 (interface value
 * 0) and 2_2_WENO5RK3_Extend/shared_utilities.cup:49-85 (interface values
 * h_fore/h_back).  p[0..6] = values at offsets -3..+3, r1/l1 = distance to
 * the crossing on the right/left. */
static void weno_nonuniform_pair(double *d_fore, double *d_back, const double p[7],
                                 double r1, double l1, double ds, double h_fore, double h_back)
{
    double x[7], h[7]; /* nodes x3m,x2m,x1m,x0,x1,x2,x3 */
    h[3] = p[3]; x[3] = 0.0;
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
    /* first divided differences d[k] over [x[k], x[k+1]] (re_step.cu:83-88) */
    double d[6];
    for (int k = 0; k < 6; ++k) d[k] = (h[k + 1] - h[k]) / (x[k + 1] - x[k]);
    /* fore: faces x2m..x3, cells d1_m1_5..d1_2_5 */
    *d_fore = weno_nonuniform_face(&x[1], &d[1], 0);
    /* back: faces x3m..x2, cells d1_m2_5..d1_1_5 */
    *d_back = weno_nonuniform_face(&x[0], &d[0], 1);
}

/* The uniform WENO5 pair of re_step.cu:260-267 (re_step.cu:11-37 on the
 * divided differences of p) in the Jiang-Peng form (G.-S. Jiang, D. Peng,
 * SIAM J. Sci. Comput. 21 (2000), eq. 2.5-2.8): with undivided second
 * differences D_j = p_j - 2 p_{j+1} + p_{j+2} and the fourth-order central
 * part C = p_1 - p_5 + 8 (p_4 - p_2),
 *   d_back = (C - Q(D0,D1,D2,D3)) / (12 ds),  d_fore = (C + Q(D4,D3,D2,D1)) / (12 ds),
 *   Q(a,b,c,d) = [4 w0 (a-2b+c) + (2 w2 - W)(b-2c+d)] / W,
 *   w0 = s1^2 s2^2, w1 = 6 s0^2 s2^2, w2 = 3 s0^2 s1^2, W = w0 + w1 + w2,
 *   s0 = (a-b)^2 + 3/13 (a-3b)^2 + e, s1 = (b-c)^2 + 3/13 (b+c)^2 + e,
 *   s2 = (c-d)^2 + 3/13 (3c-d)^2 + e,  e = 12/13 * 1e-6 * ds^2.
 * s_k = 12 ds^2/13 * (S_k + 1e-6) for the reference's smoothness indicators
 * S_k of the divided differences, so the weights are the reference's and the
 * result is algebraically the reference's; it is the exact IEEE sequence the
 * product's FP64 kernels evaluate.  The stencil-weight form
 * (or_weno_onesided, or_weno_onesided_reference) is checked against it in
 * tests/test_oracle_math.py. */
static inline double jp_q(double a, double b, double c, double d, double e)
{
    const double k = 3.0 / 13.0;
    const double t0 = a - b, t1 = b - c, t2 = c - d;
    const double y0 = fma(-3.0, b, a), y1 = b + c, y2 = fma(3.0, c, -d);
    const double s0 = fma(t0, t0, fma(k * y0, y0, e));
    const double s1 = fma(t1, t1, fma(k * y1, y1, e));
    const double s2 = fma(t2, t2, fma(k * y2, y2, e));
    const double a0 = s0 * s0, a1 = s1 * s1, a2 = s2 * s2;
    const double w0 = a1 * a2, w1 = 6.0 * (a0 * a2), w2 = 3.0 * (a0 * a1);
    const double W = (w0 + w1) + w2;
    return fma(4.0 * w0, t0 - t1, fma(2.0, w2, -W) * (t1 - t2)) / W;
}

void or_weno_pair(double *d_fore, double *d_back, const double p[7], double ds)
{
    const double k = 1.0 / (12.0 * ds), e = 12.0 / 13.0 * 1e-6 * ds * ds;
    double D[5];
    for (int j = 0; j < 5; ++j) D[j] = fma(-2.0, p[j + 1], p[j]) + p[j + 2];
    const double c = fma(8.0, p[4] - p[2], p[1] - p[5]);
    *d_back = (c - jp_q(D[0], D[1], D[2], D[3], e)) * k;
    *d_fore = (c + jp_q(D[4], D[3], D[2], D[1], e)) * k;
}

static inline void weno_uniform_pair(double *d_fore, double *d_back, const double p[7], double ds)
{
    or_weno_pair(d_fore, d_back, p, ds);
}

/* Exported for the tests: the per-axis derivative pair of re_step
 * (weno_derivative_boundary, re_step.cu:248-269; sign-change test) or of the
 * extension (shared_utilities.cup:256-278; distance test), selected by `rule`
 * (0 = reinit sign test, interface value 0; 1 = extension distance test). */
void or_weno_derivative_boundary(double *d_fore, double *d_back, const double p[7], double r1,
                                 double l1, double ds, double v_fore, double v_back, int rule)
{
    int cross = rule == 0 ? (p[2] * p[3] < 0 || p[3] * p[4] < 0) : (r1 < ds || l1 < ds);
    if (cross) {
        if (rule == 0) { v_fore = 0.0; v_back = 0.0; }
        weno_nonuniform_pair(d_fore, d_back, p, r1, l1, ds, v_fore, v_back);
    } else {
        weno_uniform_pair(d_fore, d_back, p, ds);
    }
}

/* Godunov Hamiltonian step of re_step.cu:382-392 (also
 * 1_0_ENORK2_Reinitialization/boundary_correction.cu:189-199) */
static inline double godunov_step(int inside, double xR, double xL, double yF, double yB,
                                  double zU, double zD, double dt)
{
    if (inside)
        return (sqrt(dmax2(sq(dmin2(0.0, xL)), sq(dmax2(0.0, xR)))
                     + dmax2(sq(dmin2(0.0, yB)), sq(dmax2(0.0, yF)))
                     + dmax2(sq(dmin2(0.0, zD)), sq(dmax2(0.0, zU)))) - 1) * dt * (-1.);
    return (sqrt(dmax2(sq(dmax2(0.0, xL)), sq(dmin2(0.0, xR)))
                 + dmax2(sq(dmax2(0.0, yB)), sq(dmin2(0.0, yF)))
                 + dmax2(sq(dmax2(0.0, zD)), sq(dmin2(0.0, zU)))) - 1) * dt * (1.);
}

#define SIX_DIST_ARGS const double *xpr, const double *xpl, const double *ypf, \
                      const double *ypb, const double *zpu, const double *zpd

/* Kernel re_step (WENO5): re_step.cu:283-393 */
void or_re_step(double *step, const double *lsf, const uint8_t *mask, const double *deltat,
                SIX_DIST_ARGS, int rows, int cols, int pges, double dx, double dy, double dz,
                int num_ele)
{
    (void)num_ele;
    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                double eps = 1e-6 * dx;
                if (xpr[ind] < eps || xpl[ind] < eps || ypf[ind] < eps || ypb[ind] < eps ||
                    zpu[ind] < eps || zpd[ind] < eps) {
                    step[ind] = 0;
                    continue;
                }
                double fore[3], back[3];
                for (int a = 0; a < 3; ++a) {
                    double v[7];
                    line7(lsf, r, c, p, a, rows, cols, pges, v);
                    or_weno_derivative_boundary(&fore[a], &back[a], v, R[a][ind], L[a][ind],
                                                sp[a], 0.0, 0.0, 0);
                }
                step[ind] = godunov_step(mask[ind] != 0, fore[0], back[0], fore[1], back[1],
                                         fore[2], back[2], deltat[ind]);
            }
}

/* ENO2 one-sided derivatives with the subcell node value 0:
 * 1_0_ENORK2_Reinitialization/boundary_correction.cu:80-100 */
static inline void eno2_pair(double *sR, double *sL, double v4, double v1, double v0, double v2,
                             double v3, double pr, double pl, double ds)
{
    double p2 = v1 - 2.0 * v0 + v2;
    double p2r = v0 - 2.0 * v2 + v3;
    double p2m = 0.5 * min_mod(p2, p2r) / sq(ds);
    double vr = (pr == ds) ? v2 : 0;
    *sR = (vr - v0) / pr - pr * p2m;
    double p2l = v0 - 2.0 * v1 + v4;
    p2m = 0.5 * min_mod(p2, p2l) / sq(ds);
    double vl = (pl == ds) ? v1 : 0;
    *sL = (v0 - vl) / pl + pl * p2m;
}

/* Kernel re_step (ENO2): boundary_correction.cu:147-200 (no frozen-node test) */
void or_eno2_re_step(double *step, const double *lsf, const uint8_t *mask, const double *deltat,
                     SIX_DIST_ARGS, int rows, int cols, int pges, double dx, double dy,
                     double dz, int num_ele)
{
    (void)num_ele;
    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                double fw[3], bw[3];
                for (int a = 0; a < 3; ++a) {
                    double v[7];
                    line7(lsf, r, c, p, a, rows, cols, pges, v);
                    eno2_pair(&fw[a], &bw[a], v[1], v[2], v[3], v[4], v[5], R[a][ind],
                              L[a][ind], sp[a]);
                }
                step[ind] = godunov_step(mask[ind] != 0, fw[0], bw[0], fw[1], bw[1], fw[2],
                                         bw[2], deltat[ind]);
            }
}

/* Kernel upwind_normal: 2_2_WENO5RK3_Extend/upwind_normal.cu:4-104 */
void or_upwind_normal(double *nx, double *ny, double *nz, const double *lsf, SIX_DIST_ARGS,
                      int rows, int cols, int pges, double dx, double dy, double dz, int num_ele)
{
    (void)num_ele;
    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};
    double *out[3] = {nx, ny, nz};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                for (int a = 0; a < 3; ++a) {
                    double v[7], df, db;
                    line7(lsf, r, c, p, a, rows, cols, pges, v);
                    or_weno_derivative_boundary(&df, &db, v, R[a][ind], L[a][ind], sp[a], 0.0,
                                                0.0, 1);
                    out[a][ind] = (fabs(v[4]) < fabs(v[2])) ? df : db;
                }
            }
}

/* sixth_interp: 2_2_WENO5RK3_Extend/boundary_interpolant.cu:77-99 */
static void sixth_interp(double *c_fore, double *c_back, double dis_f, double dis_b, double s,
                         const double v[7])
{
    *c_fore = 0.0;
    *c_back = 0.0;
    if (dis_f < s) {
        double xf[6] = {s, 2 * s, 3 * s, 4 * s, 5 * s, 6 * s};
        double coef[6];
        poly5_interp(xf, &v[1], coef);
        *c_fore = poly5_eval(coef, 3 * s + dis_f);
    }
    if (dis_b < s) {
        double xb[6] = {0.0, s, 2 * s, 3 * s, 4 * s, 5 * s};
        double coef[6];
        poly5_interp(xb, &v[0], coef);
        *c_back = poly5_eval(coef, 3 * s - dis_b);
    }
}

/* Kernel boundary_interpolant: boundary_interpolant.cu:102-141 */
void or_boundary_interpolant(double *cpr, double *cpl, double *cpf, double *cpb, double *cpu,
                             double *cpd, SIX_DIST_ARGS, const double *lsf, int rows, int cols,
                             int pges, double dx, double dy, double dz, int num_ele)
{
    (void)num_ele;
    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};
    double *CF[3] = {cpr, cpf, cpu}, *CB[3] = {cpl, cpb, cpd};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                for (int a = 0; a < 3; ++a) {
                    double v[7];
                    line7(lsf, r, c, p, a, rows, cols, pges, v);
                    sixth_interp(&CF[a][ind], &CB[a][ind], R[a][ind], L[a][ind], sp[a], v);
                }
            }
}

/* Kernel extend_step: 2_2_WENO5RK3_Extend/extend_step.cu:10-125 */
void or_extend_step(double *step, const double *deltat, const double *lsf, const double *vx,
                    const double *vy, const double *vz, SIX_DIST_ARGS, const double *cpr,
                    const double *cpl, const double *cpf, const double *cpb, const double *cpu,
                    const double *cpd, int rows, int cols, int pges, double dx, double dy,
                    double dz, int num_ele)
{
    (void)num_ele;
    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};
    const double *CF[3] = {cpr, cpf, cpu}, *CB[3] = {cpl, cpb, cpd};
    const double *V[3] = {vx, vy, vz};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                double eps = 1e-6 * dx;
                if (xpr[ind] < eps || xpl[ind] < eps || ypf[ind] < eps || ypb[ind] < eps ||
                    zpu[ind] < eps || zpd[ind] < eps) {
                    step[ind] = 0;
                    continue;
                }
                double acc = 0.0;
                double terms[6];
                for (int a = 0; a < 3; ++a) {
                    double v[7], df, db;
                    line7(lsf, r, c, p, a, rows, cols, pges, v);
                    or_weno_derivative_boundary(&df, &db, v, R[a][ind], L[a][ind], sp[a],
                                                CF[a][ind], CB[a][ind], 1);
                    double vel = V[a][ind];
                    terms[2 * a] = dmin2(0.0, vel) * df;
                    terms[2 * a + 1] = dmax2(0.0, vel) * db;
                }
                /* extend_step.cu:121-123 sums left to right */
                acc = terms[0] + terms[1] + terms[2] + terms[3] + terms[4] + terms[5];
                step[ind] = acc * deltat[ind];
            }
}

/* plain WENO5 fore/back derivative pair along one axis (advection_step.cu:10-28) */
static inline void plain_pair(double *df, double *db, const double *f, int r, int c, int p,
                              int axis, int rows, int cols, int pges, double ds)
{
    double v[7];
    line7(f, r, c, p, axis, rows, cols, pges, v);
    weno_uniform_pair(df, db, v, ds);
}

/* Kernel advection_step: 4_0_Calculus_ToolBox/advection_step.cu:30-56 (no dt) */
void or_advection_step(double *astep, const double *vx, const double *vy, const double *vz,
                       const double *lsf, int rows, int cols, int pges, double dx, double dy,
                       double dz)
{
    const double sp[3] = {dx, dy, dz};
    const double *V[3] = {vx, vy, vz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                double t[6];
                for (int a = 0; a < 3; ++a) {
                    double df, db;
                    plain_pair(&df, &db, lsf, r, c, p, a, rows, cols, pges, sp[a]);
                    t[2 * a] = dmin2(0.0, V[a][ind]) * df;
                    t[2 * a + 1] = dmax2(0.0, V[a][ind]) * db;
                }
                astep[ind] = t[0] + t[1] + t[2] + t[3] + t[4] + t[5];
            }
}

/* Kernel weno_derivative: 5_0_Nonoscillating_Reconstruction/nonoscillating_interpolant.cu:59-122 */
void or_weno_derivative(double *back_x, double *fore_x, double *back_y, double *fore_y,
                        double *back_z, double *fore_z, const double *lsf, int rows, int cols,
                        int pges, double dx, double dy, double dz)
{
    double *B[3] = {back_x, back_y, back_z}, *F[3] = {fore_x, fore_y, fore_z};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                for (int a = 0; a < 3; ++a)
                    plain_pair(&F[a][ind], &B[a][ind], lsf, r, c, p, a, rows, cols, pges, sp[a]);
            }
}

/* Kernel set_calculus_toolbox: 4_0_Calculus_ToolBox/set_calculus_toolbox.cu:9-107.
 * out[0..17] = Fx,Fy,Fz,FGradMag,Nx,Ny,Nz,Fxx,Fyy,Fzz,Fxy,Fyz,Fzx,FLaplacian,
 *              MeanCurvature,GaussianCurvature,Heaviside,DiracDelta */
void or_set_calculus_toolbox(double **out, const double *lsf, const double *hp, int rows,
                             int cols, int pges, double dx, double dy, double dz, double ds,
                             int num_ele)
{
    (void)num_ele;
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
#define F_(dr, dc, dp) lsf[at(r + (dr), c + (dc), p + (dp), rows, cols, pges)]
#define P_(dr, dc, dp) hp[at(r + (dr), c + (dc), p + (dp), rows, cols, pges)]
                idx_t ind = at(r, c, p, rows, cols, pges);
                double fx = (F_(0, 1, 0) - F_(0, -1, 0)) / (2 * dx);
                double fy = (F_(1, 0, 0) - F_(-1, 0, 0)) / (2 * dy);
                double fz = (F_(0, 0, 1) - F_(0, 0, -1)) / (2 * dz);
                double g = sqrt(fx * fx + fy * fy + fz * fz);
                g = dmax2(g, 1e-14);
                out[0][ind] = fx; out[1][ind] = fy; out[2][ind] = fz; out[3][ind] = g;
                out[4][ind] = fx / g; out[5][ind] = fy / g; out[6][ind] = fz / g;
                double f0 = F_(0, 0, 0);
                double fxx = (F_(0, 1, 0) - 2 * f0 + F_(0, -1, 0)) / (dx * dx);
                double fyy = (F_(1, 0, 0) - 2 * f0 + F_(-1, 0, 0)) / (dy * dy);
                double fzz = (F_(0, 0, 1) - 2 * f0 + F_(0, 0, -1)) / (dz * dz);
                double fxy = (F_(1, 1, 0) + F_(-1, -1, 0) - F_(1, -1, 0) - F_(-1, 1, 0)) / (4 * ds * ds);
                double fyz = (F_(1, 0, 1) + F_(-1, 0, -1) - F_(1, 0, -1) - F_(-1, 0, 1)) / (4 * ds * ds);
                double fzx = (F_(0, 1, 1) + F_(0, -1, -1) - F_(0, 1, -1) - F_(0, -1, 1)) / (4 * ds * ds);
                double lap = fxx + fyy + fzz;
                out[7][ind] = fxx; out[8][ind] = fyy; out[9][ind] = fzz;
                out[10][ind] = fxy; out[11][ind] = fyz; out[12][ind] = fzx; out[13][ind] = lap;
                double c1 = fxx * fx + fxy * fy + fzx * fz;
                double c2 = fxy * fx + fyy * fy + fyz * fz;
                double c3 = fzx * fx + fyz * fy + fzz * fz;
                out[14][ind] = -lap / g + (fx * c1 + fy * c2 + fz * c3) / (g * g * g);
                c1 = (fyy * fzz - fyz * fyz) * fx + (fzx * fyz - fxy * fzz) * fy + (fxy * fyz - fzx * fyy) * fz;
                c2 = (fyz * fzx - fxy * fzz) * fx + (fxx * fzz - fzx * fzx) * fy + (fzx * fxy - fxx * fyz) * fz;
                c3 = (fxy * fyz - fyy * fzx) * fx + (fzx * fxy - fxx * fyz) * fy + (fxx * fyy - fxy * fxy) * fz;
                out[15][ind] = (fx * c1 + fy * c2 + fz * c3) / (g * g * g * g);
                double px = (P_(0, 1, 0) - P_(0, -1, 0)) / (2 * dx);
                double py = (P_(1, 0, 0) - P_(-1, 0, 0)) / (2 * dy);
                double pz = (P_(0, 0, 1) - P_(0, 0, -1)) / (2 * dz);
                double dot = px * fx + py * fy + pz * fz;
                out[16][ind] = dot / (g * g);
                double p0 = P_(0, 0, 0);
                double pxx = (P_(0, 1, 0) - 2 * p0 + P_(0, -1, 0)) / (dx * dx);
                double pyy = (P_(1, 0, 0) - 2 * p0 + P_(-1, 0, 0)) / (dy * dy);
                double pzz = (P_(0, 0, 1) - 2 * p0 + P_(0, 0, -1)) / (dz * dz);
                double plap = pxx + pyy + pzz;
                out[17][ind] = plap / (g * g) - dot * lap / (g * g * g * g);
#undef F_
#undef P_
            }
}

static inline void cross3(double *w, double ux, double uy, double uz, double vx, double vy, double vz)
{
    /* cross_product, 4_0_Calculus_ToolBox/shared_utilities.cup:14-19 */
    w[0] = uy * vz - uz * vy;
    w[1] = uz * vx - ux * vz;
    w[2] = ux * vy - uy * vx;
}

/* Kernel auxi_set_calculus_toolbox: 4_0_Calculus_ToolBox/auxi_set_calculus_toolbox.cu:9-199
 * (norm = max(|v|,1e-14), shared_utilities.cup:7-11).  A is the auxiliary field
 * (lsf), ahp its primal; fin[0..12] = Fx,Fy,Fz,FGradMag,Nx,Ny,Nz,Fxx,Fyy,Fzz,Fxy,
 * Fyz,Fzx of F.  out[0..26] = Ax,Ay,Az,AGradMag,Cx,Cy,Cz,NormCrossAF,tx,ty,tz,
 * nx,ny,nz,Axx,Ayy,Azz,Axy,Ayz,Azx,ALaplacian,GeodesicCurvature,NormalCurvature,
 * GeodesicTorsion,BPerpendicular,AHeaviside,ADiracDelta.  pow(g,k) is written as
 * the product g*...*g. */
void or_auxi_set_calculus_toolbox(double **out, const double *A, const double *ahp, const double **fin,
                                  int rows, int cols, int pges, double dx, double dy, double dz, double ds,
                                  int num_ele)
{
    (void)num_ele;
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
#define A_(dr, dc, dp) A[at(r + (dr), c + (dc), p + (dp), rows, cols, pges)]
#define P_(dr, dc, dp) ahp[at(r + (dr), c + (dc), p + (dp), rows, cols, pges)]
                idx_t ind = at(r, c, p, rows, cols, pges);
                double ax = (A_(0, 1, 0) - A_(0, -1, 0)) / (2 * dx);
                double ay = (A_(1, 0, 0) - A_(-1, 0, 0)) / (2 * dy);
                double az = (A_(0, 0, 1) - A_(0, 0, -1)) / (2 * dz);
                double ag = dmax2(sqrt(ax * ax + ay * ay + az * az), 1e-14);
                out[0][ind] = ax; out[1][ind] = ay; out[2][ind] = az; out[3][ind] = ag;
                double fx = fin[0][ind], fy = fin[1][ind], fz = fin[2][ind], fg = fin[3][ind];
                double C[3];
                cross3(C, fx, fy, fz, ax, ay, az);
                double nc = dmax2(sqrt(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]), 1e-14);
                out[4][ind] = C[0]; out[5][ind] = C[1]; out[6][ind] = C[2]; out[7][ind] = nc;
                double t[3] = {C[0] / nc, C[1] / nc, C[2] / nc};
                out[8][ind] = t[0]; out[9][ind] = t[1]; out[10][ind] = t[2];
                double n[3];
                cross3(n, t[0], t[1], t[2], fin[4][ind], fin[5][ind], fin[6][ind]);
                out[11][ind] = n[0]; out[12][ind] = n[1]; out[13][ind] = n[2];
                double a0 = A_(0, 0, 0);
                double axx = (A_(0, 1, 0) - 2 * a0 + A_(0, -1, 0)) / (dx * dx);
                double ayy = (A_(1, 0, 0) - 2 * a0 + A_(-1, 0, 0)) / (dy * dy);
                double azz = (A_(0, 0, 1) - 2 * a0 + A_(0, 0, -1)) / (dz * dz);
                double axy = (A_(1, 1, 0) + A_(-1, -1, 0) - A_(1, -1, 0) - A_(-1, 1, 0)) / (4 * ds * ds);
                double ayz = (A_(1, 0, 1) + A_(-1, 0, -1) - A_(1, 0, -1) - A_(-1, 0, 1)) / (4 * ds * ds);
                double azx = (A_(0, 1, 1) + A_(0, -1, -1) - A_(0, 1, -1) - A_(0, -1, 1)) / (4 * ds * ds);
                double alap = axx + ayy + azz;
                out[14][ind] = axx; out[15][ind] = ayy; out[16][ind] = azz;
                out[17][ind] = axy; out[18][ind] = ayz; out[19][ind] = azx; out[20][ind] = alap;
                double fxx = fin[7][ind], fyy = fin[8][ind], fzz = fin[9][ind];
                double fxy = fin[10][ind], fyz = fin[11][ind], fzx = fin[12][ind];
                /* geodesic curvature, :115-131: n.((H_F t) x grad A + grad F x (H_A t))/|C| */
                double w1[3], w2[3];
                cross3(w1, t[0] * fxx + t[1] * fxy + t[2] * fzx, t[0] * fxy + t[1] * fyy + t[2] * fyz,
                       t[0] * fzx + t[1] * fyz + t[2] * fzz, ax, ay, az);
                cross3(w2, fx, fy, fz, t[0] * axx + t[1] * axy + t[2] * azx, t[0] * axy + t[1] * ayy + t[2] * ayz,
                       t[0] * azx + t[1] * ayz + t[2] * azz);
                out[21][ind] = (n[0] * (w1[0] + w2[0]) + n[1] * (w1[1] + w2[1]) + n[2] * (w1[2] + w2[2])) / nc;
                /* grad N = Hess F/|grad F| - grad F (Hess F grad F)^T/|grad F|^3, :135-145 */
                double g3 = fg * fg * fg;
                double H[3][3] = {{fxx, fxy, fzx}, {fxy, fyy, fyz}, {fzx, fyz, fzz}};
                double gF[3] = {fx, fy, fz};
                double Nm[3][3];
                for (int j = 0; j < 3; ++j) {
                    double sj = H[0][j] * fx + H[1][j] * fy + H[2][j] * fz;
                    for (int i = 0; i < 3; ++i) Nm[i][j] = H[i][j] / fg - gF[i] * sj / g3;
                }
                double v[3];
                for (int i = 0; i < 3; ++i) v[i] = Nm[i][0] * t[0] + Nm[i][1] * t[1] + Nm[i][2] * t[2];
                out[22][ind] = -(t[0] * v[0] + t[1] * v[1] + t[2] * v[2]);
                for (int i = 0; i < 3; ++i) v[i] = Nm[i][0] * n[0] + Nm[i][1] * n[1] + Nm[i][2] * n[2];
                out[23][ind] = -(t[0] * v[0] + t[1] * v[1] + t[2] * v[2]);
                out[24][ind] = -(n[0] * v[0] + n[1] * v[1] + n[2] * v[2]);
                /* Heaviside and delta of A from its primal, :166-197 */
                double px = (P_(0, 1, 0) - P_(0, -1, 0)) / (2 * dx);
                double py = (P_(1, 0, 0) - P_(-1, 0, 0)) / (2 * dy);
                double pz = (P_(0, 0, 1) - P_(0, 0, -1)) / (2 * dz);
                double dot = px * ax + py * ay + pz * az;
                out[25][ind] = dot / (ag * ag);
                double p0 = P_(0, 0, 0);
                double pxx = (P_(0, 1, 0) - 2 * p0 + P_(0, -1, 0)) / (dx * dx);
                double pyy = (P_(1, 0, 0) - 2 * p0 + P_(-1, 0, 0)) / (dy * dy);
                double pzz = (P_(0, 0, 1) - 2 * p0 + P_(0, 0, -1)) / (dz * dz);
                out[26][ind] = (pxx + pyy + pzz) / (ag * ag) - dot * alap / (ag * ag * ag * ag);
#undef A_
#undef P_
            }
}

static inline idx_t wrap_at(int r, int c, int p, int rows, int cols, int pges)
{
    r = (r + rows) % rows;
    c = (c + cols) % cols;
    p = (p + pges) % pges;
    return (idx_t)p * rows * cols + (idx_t)c * rows + r;
}

/* +SD/@SDF3/GPUAsetCalculusToolBox.m:2-64 as a whole: AHPrimal = max(A,0) (:4),
 * the kernel above (:40-54), AFDiracDelta (:57-61) and ADiracDeltaDn (:63) with
 * @GD3's Gradient/Laplacian, whose neighbours wrap around the box (circshift
 * index maps, +SD/@GD3/GD3.m:112-117; Fx :218-220, Fxx :248-250, Laplacian
 * :344-346; CrossProduct/DotProduct :352-360).  ftool[0..12] as fin above,
 * ftool[13] = Heaviside of F; out[0] = AHPrimal, out[1..27] the kernel's,
 * out[28] = AFDiracDelta, out[29] = ADiracDeltaDn. */
void or_aux_calculus_toolbox(double **out, const double *A, const double **ftool, int rows, int cols,
                             int pges, double dx, double dy, double dz, double ds)
{
    idx_t n = (idx_t)rows * cols * pges;
    for (idx_t i = 0; i < n; ++i) out[0][i] = A[i] > 0 ? A[i] : 0.0; /* MATLAB max ignores NaN */
    or_auxi_set_calculus_toolbox(out + 1, A, out[0], ftool, rows, cols, pges, dx, dy, dz, ds, (int)n);
    const double *AH = out[26], *Hf = ftool[13], *AG = out[4], *NC = out[8], *ADD = out[27];
    const double *Cx = out[5], *Cy = out[6], *Cz = out[7];
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = (idx_t)p * rows * cols + (idx_t)c * rows + r;
                idx_t xp = wrap_at(r, c + 1, p, rows, cols, pges), xm = wrap_at(r, c - 1, p, rows, cols, pges);
                idx_t yp = wrap_at(r + 1, c, p, rows, cols, pges), ym = wrap_at(r - 1, c, p, rows, cols, pges);
                idx_t zp = wrap_at(r, c, p + 1, rows, cols, pges), zm = wrap_at(r, c, p - 1, rows, cols, pges);
                double h1[3] = {(AH[xp] - AH[xm]) / (2 * dx), (AH[yp] - AH[ym]) / (2 * dy),
                                (AH[zp] - AH[zm]) / (2 * dz)};
                double h2[3] = {(Hf[xp] - Hf[xm]) / (2 * dx), (Hf[yp] - Hf[ym]) / (2 * dy),
                                (Hf[zp] - Hf[zm]) / (2 * dz)};
                double ch[3];
                cross3(ch, h2[0], h2[1], h2[2], h1[0], h1[1], h1[2]);
                out[28][ind] = (Cx[ind] * ch[0] + Cy[ind] * ch[1] + Cz[ind] * ch[2]) / (NC[ind] * NC[ind]);
                double lah = (AH[xp] - 2 * AH[ind] + AH[xm]) / (dx * dx) + (AH[yp] - 2 * AH[ind] + AH[ym]) / (dy * dy)
                             + (AH[zp] - 2 * AH[ind] + AH[zm]) / (dz * dz);
                double la = (A[xp] - 2 * A[ind] + A[xm]) / (dx * dx) + (A[yp] - 2 * A[ind] + A[ym]) / (dy * dy)
                            + (A[zp] - 2 * A[ind] + A[zm]) / (dz * dz);
                out[29][ind] = (lah - ADD[ind] * la) / AG[ind];
            }
}

/* ------------------------------------------------------------------------- */
/* reductions                                                                 */
/* ------------------------------------------------------------------------- */

/* Neumaier compensated sum in index order (the reference's gpuArray sum has
 * an unspecified order; the product's GPU reduction is compared at a
 * tolerance, see tests). */
typedef struct { double s, c; } ksum_t;
static inline void kadd(ksum_t *k, double x)
{
    double t = k->s + x;
    if (fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x;
    else k->c += (x - t) + k->s;
    k->s = t;
}

/* calArea / calVolume with the Towers delta/H of set_calculus_toolbox:
 * +SD/@SDF3/SDF3.m:73-80 (masks |F|<2Ds and F<2Ds, weight Ds^3);
 * delta/H per set_calculus_toolbox.cu:91-106 with HPrimal = max(F,0)
 * (+SD/@SDF3/GPUsetCalculusToolBox.m:4). */
void or_area_volume_towers(const double *F, int rows, int cols, int pges, double dx, double dy,
                           double dz, double *area, double *volume)
{
    idx_t n = (idx_t)rows * cols * pges;
    double *hp = (double *)calloc((size_t)n, sizeof(double));
    double *buf = (double *)malloc(sizeof(double) * n * 18);
    double *out[18];
    for (int k = 0; k < 18; ++k) out[k] = buf + (idx_t)k * n;
    for (idx_t i = 0; i < n; ++i) hp[i] = dmax2(F[i], 0);
    double ds = (dx + dy + dz) / 3;
    or_set_calculus_toolbox(out, F, hp, rows, cols, pges, dx, dy, dz, ds, (int)n);
    ksum_t sa = {0, 0}, sv = {0, 0};
    for (idx_t i = 0; i < n; ++i) {
        if (fabs(F[i]) < 2 * ds) kadd(&sa, out[17][i]);
        if (F[i] < 2 * ds) kadd(&sv, 1 - out[16][i]);
    }
    *area = (sa.s + sa.c) * (ds * ds * ds);
    *volume = (sv.s + sv.c) * (ds * ds * ds);
    free(buf);
    free(hp);
}

/* Cosine-smoothed delta/H of width eps (src/0_0_setup/+SD/@SDF3/SDF3.m:126-136)
 * with the periodic central-difference gradient of src/0_0_setup/+SD/@GD3/GD3.m:212-223,
 * integrated as SurfaceIntegral(1)/VolumeIntegral(1) (SDF3.m:178-187) with
 * EltVol = Dx*Dy*Dz (GD3.m:98). */
void or_area_volume_cosine(const double *F, int rows, int cols, int pges, double dx, double dy,
                           double dz, double eps, double *area, double *volume)
{
    ksum_t sa = {0, 0}, sv = {0, 0};
    const double pi = 3.14159265358979323846;
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = (idx_t)p * rows * cols + (idx_t)c * rows + r;
                double v = F[ind];
                double delta = 0.0, heav;
                if (v > eps) {
                    heav = 1.0;
                } else if (v < -eps) {
                    heav = 0.0;
                } else {
                    int cp = (c + 1) % cols, cm = (c - 1 + cols) % cols;
                    int rp = (r + 1) % rows, rm = (r - 1 + rows) % rows;
                    int pp = (p + 1) % pges, pm = (p - 1 + pges) % pges;
#define IX(rr, cc, pq) ((idx_t)(pq) * rows * cols + (idx_t)(cc) * rows + (rr))
                    double fx = (F[IX(r, cp, p)] - F[IX(r, cm, p)]) / (2 * dx);
                    double fy = (F[IX(rp, c, p)] - F[IX(rm, c, p)]) / (2 * dy);
                    double fz = (F[IX(r, c, pp)] - F[IX(r, c, pm)]) / (2 * dz);
#undef IX
                    double g = sqrt(sq(fx) + sq(fy) + sq(fz));
                    delta = g * (1 + cos(pi * v / eps)) / eps / 2;
                    heav = (1 + v / eps + sin(pi * v / eps) / pi) / 2;
                }
                kadd(&sa, delta);
                kadd(&sv, 1 - heav);
            }
    double ev = dx * dy * dz;
    *area = (sa.s + sa.c) * ev;
    *volume = (sv.s + sv.c) * ev;
}

/* ------------------------------------------------------------------------- */
/* wrappers (MATLAB +SD/@SDF3 methods)                                        */
/* ------------------------------------------------------------------------- */

static double *dalloc(idx_t n) { return (double *)malloc(sizeof(double) * (size_t)n); }

/* setDistance: +SD/@SDF3/setDistance.m:5-21.  order 6 = boundary_location,
 * order 2 = ENO2 boundary_correction.  Writes the six arrays and min_dist. */
void or_set_distance(int order, const double *F, double *xpr, double *xpl, double *ypf,
                     double *ypb, double *zpu, double *zpd, double *min_dist, int rows, int cols,
                     int pges, double dx, double dy, double dz)
{
    idx_t n = (idx_t)rows * cols * pges;
    for (idx_t i = 0; i < n; ++i) {
        xpr[i] = dx; xpl[i] = dx; ypf[i] = dy; ypb[i] = dy; zpu[i] = dz; zpd[i] = dz;
    }
    distances_impl(order, xpr, xpl, ypf, ypb, zpu, zpd, F, rows, cols, pges, dx, dy, dz);
    if (min_dist)
        for (idx_t i = 0; i < n; ++i)
            min_dist[i] = dmin2(dmin2(xpr[i], xpl[i]), dmin2(dmin2(ypf[i], ypb[i]), dmin2(zpu[i], zpd[i])));
}

/* Shu-Osher TVD-RK3 combinations exactly as the MATLAB wrappers write them
 * (+SD/@SDF3/WENO5RK3Reinitialization.m:20-37). */
static void rk3_combine_stage2(double *out, const double *fn, const double *f1, const double *step, idx_t n)
{
    for (idx_t i = 0; i < n; ++i) {
        double t2 = f1[i] - step[i];
        out[i] = 3. / 4. * fn[i] + 1. / 4. * t2;
    }
}
static void rk3_combine_stage3(double *fn, const double *f05, const double *step, idx_t n)
{
    for (idx_t i = 0; i < n; ++i) {
        double t15 = f05[i] - step[i];
        fn[i] = 1. / 3. * fn[i] + 2. / 3. * t15;
    }
}
static void sub_into(double *out, const double *a, const double *b, idx_t n)
{
    for (idx_t i = 0; i < n; ++i) out[i] = a[i] - b[i];
}

/* WENO5RK3Reinitialization: +SD/@SDF3/WENO5RK3Reinitialization.m:3-43.
 * F is updated in place; mask = F<0 at entry; deltat = 0.3*min_dist where the
 * distances are the ones passed (from the last setDistance). */
void or_weno5_rk3_reinit(double *F, int iterations, SIX_DIST_ARGS, const double *min_dist,
                         int rows, int cols, int pges, double dx, double dy, double dz)
{
    idx_t n = (idx_t)rows * cols * pges;
    uint8_t *mask = (uint8_t *)malloc((size_t)n);
    double *dt = dalloc(n), *step = dalloc(n), *f1 = dalloc(n), *f2 = dalloc(n);
    for (idx_t i = 0; i < n; ++i) { mask[i] = F[i] < 0; dt[i] = 0.3 * min_dist[i]; }
    for (int it = 0; it < iterations; ++it) {
        or_re_step(step, F, mask, dt, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, (int)n);
        sub_into(f1, F, step, n);
        or_re_step(step, f1, mask, dt, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, (int)n);
        rk3_combine_stage2(f2, F, f1, step, n);
        or_re_step(step, f2, mask, dt, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, (int)n);
        rk3_combine_stage3(F, f2, step, n);
    }
    free(mask); free(dt); free(step); free(f1); free(f2);
}

/* ENORK2Reinitialization: +SD/@SDF3/ENORK2Reinitialization.m:3-45 with its
 * integrator exposed: rk2 != 0 is the reference's Heun RK2 (:29-40); rk2 == 0
 * is forward Euler (F <- F - step).  deltat = cfl*min of the ENO2 distances,
 * computed from the argument F (:17-25; the reference uses cfl = 0.3). */
void or_eno2_reinit(double *F, int iterations, int rk2, double cfl, int rows, int cols, int pges,
                    double dx, double dy, double dz)
{
    idx_t n = (idx_t)rows * cols * pges;
    uint8_t *mask = (uint8_t *)malloc((size_t)n);
    double *d6 = dalloc(6 * n), *md = dalloc(n), *dt = dalloc(n), *step = dalloc(n), *ft = dalloc(n);
    double *xpr = d6, *xpl = d6 + n, *ypf = d6 + 2 * n, *ypb = d6 + 3 * n, *zpu = d6 + 4 * n, *zpd = d6 + 5 * n;
    for (idx_t i = 0; i < n; ++i) mask[i] = F[i] < 0;
    or_set_distance(2, F, xpr, xpl, ypf, ypb, zpu, zpd, md, rows, cols, pges, dx, dy, dz);
    for (idx_t i = 0; i < n; ++i) dt[i] = cfl * md[i];
    for (int it = 0; it < iterations; ++it) {
        or_eno2_re_step(step, F, mask, dt, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, (int)n);
        if (!rk2) {
            for (idx_t i = 0; i < n; ++i) F[i] = F[i] - step[i];
            continue;
        }
        sub_into(ft, F, step, n);
        or_eno2_re_step(step, ft, mask, dt, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, (int)n);
        for (idx_t i = 0; i < n; ++i) F[i] = (F[i] + ft[i] - step[i]) / 2;
    }
    free(mask); free(d6); free(md); free(dt); free(step); free(ft);
}

/* Velocity of the extension: v = sign(F) * grad/max(|grad|,floor), grad from
 * upwind_normal (+SD/@SDF3/WENO5RK3Extend.m:18-31, floor 1e-6;
 * WENO5RK3ClosetPointSurfaceRedistance.m:54-74, floor 1e-14). */
void or_extension_velocity(const double *F, SIX_DIST_ARGS, double floor_, double *vx, double *vy,
                           double *vz, int rows, int cols, int pges, double dx, double dy, double dz)
{
    idx_t n = (idx_t)rows * cols * pges;
    or_upwind_normal(vx, vy, vz, F, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, (int)n);
    for (idx_t i = 0; i < n; ++i) {
        double g = dmax2(sqrt(sq(vx[i]) + sq(vy[i]) + sq(vz[i])), floor_);
        double s = msign(F[i]);
        vx[i] = s * (vx[i] / g);
        vy[i] = s * (vy[i] / g);
        vz[i] = s * (vz[i] / g);
    }
}

/* The RK3 loop shared by WENO5RK3Extend (+SD/@SDF3/WENO5RK3Extend.m:41-74)
 * and the local extension of WENO5RK3ClosetPointSurfaceRedistance.m:119-157:
 * boundary_interpolant once, then `iterations` RK3 steps of extend_step. */
void or_extend_loop(double *C, int iterations, const double *vx, const double *vy,
                    const double *vz, const double *deltat, SIX_DIST_ARGS, int rows, int cols,
                    int pges, double dx, double dy, double dz)
{
    idx_t n = (idx_t)rows * cols * pges;
    double *cp = dalloc(6 * n), *step = dalloc(n), *c1 = dalloc(n), *c2 = dalloc(n);
    or_boundary_interpolant(cp, cp + n, cp + 2 * n, cp + 3 * n, cp + 4 * n, cp + 5 * n, xpr, xpl,
                            ypf, ypb, zpu, zpd, C, rows, cols, pges, dx, dy, dz, (int)n);
#define EXT(src) or_extend_step(step, deltat, src, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cp, \
                                cp + n, cp + 2 * n, cp + 3 * n, cp + 4 * n, cp + 5 * n, rows,     \
                                cols, pges, dx, dy, dz, (int)n)
    for (int it = 0; it < iterations; ++it) {
        EXT(C);
        sub_into(c1, C, step, n);
        EXT(c1);
        rk3_combine_stage2(c2, C, c1, step, n);
        EXT(c2);
        rk3_combine_stage3(C, c2, step, n);
    }
#undef EXT
    free(cp); free(step); free(c1); free(c2);
}

/* WENO5RK3Extend: +SD/@SDF3/WENO5RK3Extend.m:9-76.  F and its distances
 * (setDistance of F) define the interface; C is extended in place. */
void or_weno5_rk3_extend(double *C, int iterations, const double *F, SIX_DIST_ARGS,
                         const double *min_dist, int rows, int cols, int pges, double dx,
                         double dy, double dz)
{
    idx_t n = (idx_t)rows * cols * pges;
    double *v = dalloc(3 * n), *dt = dalloc(n);
    or_extension_velocity(F, xpr, xpl, ypf, ypb, zpu, zpd, 1e-6, v, v + n, v + 2 * n, rows, cols,
                          pges, dx, dy, dz);
    for (idx_t i = 0; i < n; ++i) dt[i] = 0.3 * min_dist[i];
    or_extend_loop(C, iterations, v, v + n, v + 2 * n, dt, xpr, xpl, ypf, ypb, zpu, zpd, rows,
                   cols, pges, dx, dy, dz);
    free(v); free(dt);
}

/* WENO5RK3ClosetPointSurfaceRedistance: +SD/@SDF3/WENO5RK3ClosetPointSurfaceRedistance.m:4-117 */
void or_surface_redistance(double *A, const double *F, int iter1, int iter2, int rows, int cols,
                           int pges, double dx, double dy, double dz)
{
    idx_t n = (idx_t)rows * cols * pges;
    double *ad = dalloc(6 * n), *amd = dalloc(n), *adt = dalloc(n);
    double *fd = dalloc(6 * n), *fmd = dalloc(n), *fdt = dalloc(n);
    double *v = dalloc(3 * n), *step = dalloc(n), *a1 = dalloc(n), *a05 = dalloc(n);
    uint8_t *mask = (uint8_t *)malloc((size_t)n);
#define D6(b) b, b + n, b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n
    or_set_distance(6, A, D6(ad), amd, rows, cols, pges, dx, dy, dz);
    for (idx_t i = 0; i < n; ++i) { adt[i] = 0.3 * amd[i]; mask[i] = A[i] < 0; }
    or_set_distance(6, F, D6(fd), fmd, rows, cols, pges, dx, dy, dz);
    for (idx_t i = 0; i < n; ++i) fdt[i] = 0.3 * fmd[i];
    or_extension_velocity(F, D6(fd), 1e-14, v, v + n, v + 2 * n, rows, cols, pges, dx, dy, dz);
#define XT(buf) or_extend_loop(buf, iter2, v, v + n, v + 2 * n, fdt, D6(fd), rows, cols, pges, dx, dy, dz)
#define RS(src) or_re_step(step, src, mask, adt, D6(ad), rows, cols, pges, dx, dy, dz, (int)n)
    for (int it = 0; it < iter1; ++it) {
        XT(A);
        RS(A);
        sub_into(a1, A, step, n);
        XT(a1);
        RS(a1);
        rk3_combine_stage2(a05, A, a1, step, n);
        XT(a05);
        RS(a05);
        rk3_combine_stage3(A, a05, step, n);
    }
    XT(A);
#undef XT
#undef RS
#undef D6
    free(ad); free(amd); free(adt); free(fd); free(fmd); free(fdt);
    free(v); free(step); free(a1); free(a05); free(mask);
}

/* TVD-RK3 WENO5 advection with a global dt: the Shu-Osher form of
 * +SD/@SDF3/WENORK3SurfaceConservationLaw.m:10-29 with L = advection_step
 * (4_0_Calculus_ToolBox/advection_step.cu:30-56) and step = dt*L. */
void or_advect_rk3(double *F, int steps, double dt, const double *vx, const double *vy,
                   const double *vz, int rows, int cols, int pges, double dx, double dy, double dz)
{
    idx_t n = (idx_t)rows * cols * pges;
    double *a = dalloc(n), *step = dalloc(n), *f1 = dalloc(n), *f2 = dalloc(n);
    for (int s = 0; s < steps; ++s) {
        or_advection_step(a, vx, vy, vz, F, rows, cols, pges, dx, dy, dz);
        for (idx_t i = 0; i < n; ++i) step[i] = dt * a[i];
        sub_into(f1, F, step, n);
        or_advection_step(a, vx, vy, vz, f1, rows, cols, pges, dx, dy, dz);
        for (idx_t i = 0; i < n; ++i) step[i] = dt * a[i];
        rk3_combine_stage2(f2, F, f1, step, n);
        or_advection_step(a, vx, vy, vz, f2, rows, cols, pges, dx, dy, dz);
        for (idx_t i = 0; i < n; ++i) step[i] = dt * a[i];
        rk3_combine_stage3(F, f2, step, n);
    }
    free(a); free(step); free(f1); free(f2);
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* u.N with N = grad F / max(|grad F|, 1e-14) from clamped central differences
 * (the normal of set_calculus_toolbox.cu:31-44); the field the drivers extend
 * off the interface (p00_test_singlephase_relax.m:76-84). */
void or_normal_velocity(double *out, const double *F, const double *vx, const double *vy, const double *vz,
                        int rows, int cols, int pges, double dx, double dy, double dz)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                double fx = (F[at(r, c + 1, p, rows, cols, pges)] - F[at(r, c - 1, p, rows, cols, pges)]) / (2 * dx);
                double fy = (F[at(r + 1, c, p, rows, cols, pges)] - F[at(r - 1, c, p, rows, cols, pges)]) / (2 * dy);
                double fz = (F[at(r, c, p + 1, rows, cols, pges)] - F[at(r, c, p - 1, rows, cols, pges)]) / (2 * dz);
                double g = dmax2(sqrt(fx * fx + fy * fy + fz * fz), 1e-14);
                out[ind] = vx[ind] * (fx / g) + vy[ind] * (fy / g) + vz[ind] * (fz / g);
            }
}

/* ------------------------------------------------------------------------- */
/* WENO-RK3 / ENO-RK2 variants of the reinitialization and the extension     */
/* (CUDA_Code/1_1_WENORK3_Reinitialization, 2_0_ENORK2_Extend,               */
/* 2_1_WENORK3_Extend; +SD/@SDF3/WENORK3Reinitialization.m, ENORK2Extend.m,   */
/* WENORK3Extend.m)                                                          */
/* ------------------------------------------------------------------------- */
static inline double min_abs(double x, double y) { return (fabs(x) < fabs(y)) ? x : y; }

/* select_stencil: weno_rk3_reinitialization.cu:149-184 (interface value 0),
 * wenork3_extend.cu:88-123 (interface values h_fore/h_back).  x[0..6] =
 * x3m..x3, h[0..6] the values; p[0..6] = p1..p7. */
static void cubic_stencil(double x[7], double h[7], const double p[7], double r1, double l1, double ds,
                          double h_fore, double h_back)
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

/* ENO_cubic_derivative: weno_rk3_reinitialization.cu:190-222 (the same text in
 * wenork3_extend.cu:129-164).  Exported for the tests. */
void or_eno_cubic_pair(double *d_fore, double *d_back, const double x[7], const double h[7])
{
    double d1_25 = (h[6] - h[5]) / (x[6] - x[5]);
    double d1_15 = (h[5] - h[4]) / (x[5] - x[4]);
    double d1_05 = (h[4] - h[3]) / (x[4] - x[3]);
    double d1_m05 = (h[3] - h[2]) / (x[3] - x[2]);
    double d1_m15 = (h[2] - h[1]) / (x[2] - x[1]);
    double d1_m25 = (h[1] - h[0]) / (x[1] - x[0]);
    double d2_2 = (d1_25 - d1_15) / (x[6] - x[4]);
    double d2_1 = (d1_15 - d1_05) / (x[5] - x[3]);
    double d2_0 = (d1_05 - d1_m05) / (x[4] - x[2]);
    double d2_m1 = (d1_m05 - d1_m15) / (x[3] - x[1]);
    double d2_m2 = (d1_m15 - d1_m25) / (x[2] - x[0]);
    double d3_15 = (d2_2 - d2_1) / (x[6] - x[3]);
    double d3_05 = (d2_1 - d2_0) / (x[5] - x[2]);
    double d3_m05 = (d2_0 - d2_m1) / (x[4] - x[1]);
    double d3_m15 = (d2_m1 - d2_m2) / (x[3] - x[0]);
    double a1 = (x[3] - x[2]) * (x[3] - x[1]) * min_abs(d3_m05, d3_m15);
    double a2 = (x[3] - x[2]) * (x[3] - x[4]) * min_abs(d3_m05, d3_05);
    double a = (fabs(d2_m1) < fabs(d2_0)) ? a1 : a2;
    double b1 = (x[3] - x[2]) * (x[3] - x[4]) * min_abs(d3_m05, d3_05);
    double b2 = (x[3] - x[4]) * (x[3] - x[5]) * min_abs(d3_05, d3_15);
    double b = (fabs(d2_0) < fabs(d2_1)) ? b1 : b2;
    *d_back = d1_m05 + min_mod(d2_m1, d2_0) * (x[3] - x[2]) + a;
    *d_fore = d1_05 + min_mod(d2_0, d2_1) * (x[3] - x[4]) + b;
}

/* weno_derivative_boundary of the variants: rule 0 = the reinitialization's
 * sign-change test with interface value 0 (weno_rk3_reinitialization.cu:228-247),
 * rule 1 = the extension's subcell-distance test with the interface values
 * v_fore/v_back (wenork3_extend.cu:167-190).  The uniform branch is the WENO5
 * pair of re_step.cu:11-37 (or_weno_pair). */
void or_wenork3_derivative_boundary(double *d_fore, double *d_back, const double p[7], double r1,
                                    double l1, double ds, double v_fore, double v_back, int rule)
{
    int cross = rule == 0 ? (p[2] * p[3] < 0 || p[3] * p[4] < 0) : (r1 < ds || l1 < ds);
    if (cross) {
        double x[7], h[7];
        if (rule == 0) { v_fore = 0.0; v_back = 0.0; }
        cubic_stencil(x, h, p, r1, l1, ds, v_fore, v_back);
        or_eno_cubic_pair(d_fore, d_back, x, h);
    } else {
        weno_uniform_pair(d_fore, d_back, p, ds);
    }
}

/* Kernel re_step (WENO-RK3): weno_rk3_reinitialization.cu:295-407 */
void or_wenork3_re_step(double *step, const double *lsf, const uint8_t *mask, const double *deltat,
                        SIX_DIST_ARGS, int rows, int cols, int pges, double dx, double dy, double dz,
                        int num_ele)
{
    (void)num_ele;
    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                double eps = 1e-6 * dx;
                if (xpr[ind] < eps || xpl[ind] < eps || ypf[ind] < eps || ypb[ind] < eps ||
                    zpu[ind] < eps || zpd[ind] < eps) {
                    step[ind] = 0;
                    continue;
                }
                double fore[3], back[3];
                for (int a = 0; a < 3; ++a) {
                    double v[7];
                    line7(lsf, r, c, p, a, rows, cols, pges, v);
                    or_wenork3_derivative_boundary(&fore[a], &back[a], v, R[a][ind], L[a][ind], sp[a],
                                                   0.0, 0.0, 0);
                }
                step[ind] = godunov_step(mask[ind] != 0, fore[0], back[0], fore[1], back[1],
                                         fore[2], back[2], deltat[ind]);
            }
}

/* the upwind combination of the extension steps (enork2_extend.cu:262-264,
 * wenork3_extend.cu:490-492), summed left to right */
static inline double upwind_sum(const double fo[3], const double bk[3], double vx, double vy, double vz,
                                double dt)
{
    return (dmin2(0.0, vx) * fo[0] + dmax2(0.0, vx) * bk[0] + dmin2(0.0, vy) * fo[1] +
            dmax2(0.0, vy) * bk[1] + dmin2(0.0, vz) * fo[2] + dmax2(0.0, vz) * bk[2]) * dt;
}

/* eno_derivative_field: enork2_extend.cu:56-78 */
static void eno2_field_pair(double *sR, double *sL, const double p[7], double dis_f, double dis_b,
                            double ds, double v_fore, double v_back)
{
    double v4 = p[1], v1 = p[2], v0 = p[3], v2 = p[4], v3 = p[5];
    double p2 = v1 - 2.0 * v0 + v2;
    double p2r = v0 - 2.0 * v2 + v3;
    double p2m = 0.5 * min_mod(p2, p2r) / (ds * ds);
    double v_f = (dis_f == ds) ? v2 : v_fore;
    *sR = (v_f - v0) / dis_f - dis_f * p2m;
    double p2l = v0 - 2.0 * v1 + v4;
    p2m = 0.5 * min_mod(p2, p2l) / (ds * ds);
    double v_b = (dis_b == ds) ? v1 : v_back;
    *sL = (v0 - v_b) / dis_b + dis_b * p2m;
}

/* Kernel upwind_normal (ENO-RK2): enork2_extend.cu:102-138 with
 * upwind_normal_point :82-99 */
void or_enork2_upwind_normal(double *nx, double *ny, double *nz, const double *lsf, SIX_DIST_ARGS,
                             int rows, int cols, int pges, double dx, double dy, double dz, int num_ele)
{
    (void)num_ele;
    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};
    double *out[3] = {nx, ny, nz};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                for (int a = 0; a < 3; ++a) {
                    double v[7], sR, sL;
                    line7(lsf, r, c, p, a, rows, cols, pges, v);
                    double pr = R[a][ind], pl = L[a][ind];
                    eno2_pair(&sR, &sL, v[1], v[2], v[3], v[4], v[5], pr, pl, sp[a]);
                    double vr = (pr == sp[a]) ? v[4] : 0, vl = (pl == sp[a]) ? v[2] : 0;
                    out[a][ind] = (fabs(vr) < fabs(vl)) ? sR : sL;
                }
            }
}

/* cubic_eno_interp: enork2_extend.cu:144-170 */
static void eno2_interface_values(double *c_fore, double *c_back, double dis_f, double dis_b, double ds,
                                  const double p[7])
{
    double v4 = p[1], v1 = p[2], v0 = p[3], v2 = p[4], v3 = p[5];
    *c_fore = 0;
    *c_back = 0;
    if (dis_f != ds) {
        double p2 = v2 - 2.0 * v0 + v1;
        double p2r = v3 - 2.0 * v2 + v0;
        double c2 = 0.5 * min_mod(p2, p2r) / (ds * ds);
        double c1 = (v2 - v0) / ds;
        double c0 = (v2 + v0) / 2 - c2 * (ds * ds) / 4;
        double x = dis_f - ds / 2;
        *c_fore = c2 * x * x + c1 * x + c0;
    }
    if (dis_b != ds) {
        double p2 = v2 - 2.0 * v0 + v1;
        double p2l = v0 - 2.0 * v1 + v4;
        double c2 = 0.5 * min_mod(p2, p2l) / (ds * ds);
        double c1 = (v0 - v1) / ds;
        double c0 = (v1 + v0) / 2 - c2 * (ds * ds) / 4;
        double x = ds / 2 - dis_b;
        *c_back = c2 * x * x + c1 * x + c0;
    }
}

/* cubic_interp_coefficient (wenork3_extend.cu:297-306) and cubic_interp
 * (:310-329); pow(xc,2), pow(xc,3) are xc*xc and xc*xc*xc */
static void cubic_coefficients(double c[4], double v0, double v1, double v2, double v3, double s)
{
    c[0] = v0;
    c[1] = (3.0 * (v1 - v0) - 3.0 / 2.0 * (v2 - v0) + 1.0 / 3.0 * (v3 - v0)) / s;
    c[2] = (-5.0 / 2.0 * (v1 - v0) + 2.0 * (v2 - v0) - 1.0 / 2.0 * (v3 - v0)) / (s * s);
    c[3] = (1.0 / 2.0 * (v1 - v0) - 1.0 / 2.0 * (v2 - v0) + 1.0 / 6.0 * (v3 - v0)) / (s * s * s);
}
static void cubic_interface_values(double *c_fore, double *c_back, double dis_f, double dis_b, double ds,
                                   const double p[7])
{
    double c[4];
    *c_fore = 0.;
    *c_back = 0.;
    if (dis_f < ds) {
        cubic_coefficients(c, p[2], p[3], p[4], p[5], ds);
        double xc = ds + dis_f;
        *c_fore = c[0] + c[1] * xc + c[2] * (xc * xc) + c[3] * (xc * xc * xc);
    }
    if (dis_b < ds) {
        cubic_coefficients(c, p[1], p[2], p[3], p[4], ds);
        double xc = 2 * ds - dis_b;
        *c_back = c[0] + c[1] * xc + c[2] * (xc * xc) + c[3] * (xc * xc * xc);
    }
}

/* Kernels boundary_interpolate: enork2_extend.cu:174-210 (weno == 0) and
 * wenork3_extend.cu:331-368 (weno == 1) */
static void variant_interpolate(int weno, double *cpr, double *cpl, double *cpf, double *cpb, double *cpu,
                                double *cpd, SIX_DIST_ARGS, const double *lsf, int rows, int cols, int pges,
                                double dx, double dy, double dz)
{
    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};
    double *CF[3] = {cpr, cpf, cpu}, *CB[3] = {cpl, cpb, cpd};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                for (int a = 0; a < 3; ++a) {
                    double v[7];
                    line7(lsf, r, c, p, a, rows, cols, pges, v);
                    if (weno) cubic_interface_values(&CF[a][ind], &CB[a][ind], R[a][ind], L[a][ind], sp[a], v);
                    else eno2_interface_values(&CF[a][ind], &CB[a][ind], R[a][ind], L[a][ind], sp[a], v);
                }
            }
}

void or_enork2_boundary_interpolate(double *cpr, double *cpl, double *cpf, double *cpb, double *cpu,
                                    double *cpd, SIX_DIST_ARGS, const double *lsf, int rows, int cols,
                                    int pges, double dx, double dy, double dz, int num_ele)
{
    (void)num_ele;
    variant_interpolate(0, cpr, cpl, cpf, cpb, cpu, cpd, xpr, xpl, ypf, ypb, zpu, zpd, lsf, rows, cols,
                        pges, dx, dy, dz);
}

void or_wenork3_boundary_interpolate(double *cpr, double *cpl, double *cpf, double *cpb, double *cpu,
                                     double *cpd, SIX_DIST_ARGS, const double *lsf, int rows, int cols,
                                     int pges, double dx, double dy, double dz, int num_ele)
{
    (void)num_ele;
    variant_interpolate(1, cpr, cpl, cpf, cpb, cpu, cpd, xpr, xpl, ypf, ypb, zpu, zpd, lsf, rows, cols,
                        pges, dx, dy, dz);
}

/* Kernels extend_step: enork2_extend.cu:213-266 (weno == 0, no frozen test)
 * and wenork3_extend.cu:372-495 (weno == 1, frozen below 1e-6 dx) */
static void variant_extend_step(int weno, double *step, const double *deltat, const double *lsf,
                                const double *vx, const double *vy, const double *vz, SIX_DIST_ARGS,
                                const double *cpr, const double *cpl, const double *cpf, const double *cpb,
                                const double *cpu, const double *cpd, int rows, int cols, int pges, double dx,
                                double dy, double dz)
{
    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};
    const double *CF[3] = {cpr, cpf, cpu}, *CB[3] = {cpl, cpb, cpd};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                double eps = 1e-6 * dx;
                if (weno && (xpr[ind] < eps || xpl[ind] < eps || ypf[ind] < eps || ypb[ind] < eps ||
                             zpu[ind] < eps || zpd[ind] < eps)) {
                    step[ind] = 0;
                    continue;
                }
                double fo[3], bk[3];
                for (int a = 0; a < 3; ++a) {
                    double v[7];
                    line7(lsf, r, c, p, a, rows, cols, pges, v);
                    if (weno)
                        or_wenork3_derivative_boundary(&fo[a], &bk[a], v, R[a][ind], L[a][ind], sp[a],
                                                       CF[a][ind], CB[a][ind], 1);
                    else
                        eno2_field_pair(&fo[a], &bk[a], v, R[a][ind], L[a][ind], sp[a], CF[a][ind],
                                        CB[a][ind]);
                }
                step[ind] = upwind_sum(fo, bk, vx[ind], vy[ind], vz[ind], deltat[ind]);
            }
}

void or_enork2_extend_step(double *step, const double *deltat, const double *lsf, const double *vx,
                           const double *vy, const double *vz, SIX_DIST_ARGS, const double *cpr,
                           const double *cpl, const double *cpf, const double *cpb, const double *cpu,
                           const double *cpd, int rows, int cols, int pges, double dx, double dy,
                           double dz, int num_ele)
{
    (void)num_ele;
    variant_extend_step(0, step, deltat, lsf, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cpr, cpl, cpf, cpb,
                        cpu, cpd, rows, cols, pges, dx, dy, dz);
}

void or_wenork3_extend_step(double *step, const double *deltat, const double *lsf, const double *vx,
                            const double *vy, const double *vz, SIX_DIST_ARGS, const double *cpr,
                            const double *cpl, const double *cpf, const double *cpb, const double *cpu,
                            const double *cpd, int rows, int cols, int pges, double dx, double dy,
                            double dz, int num_ele)
{
    (void)num_ele;
    variant_extend_step(1, step, deltat, lsf, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cpr, cpl, cpf, cpb,
                        cpu, cpd, rows, cols, pges, dx, dy, dz);
}

/* Kernel upwind_normal (WENO-RK3): wenork3_extend.cu:202-292 with
 * upwind_normal_point :193-199 (interface values 0, distance test) */
void or_wenork3_upwind_normal(double *nx, double *ny, double *nz, const double *lsf, SIX_DIST_ARGS,
                              int rows, int cols, int pges, double dx, double dy, double dz, int num_ele)
{
    (void)num_ele;
    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};
    double *out[3] = {nx, ny, nz};
    const double sp[3] = {dx, dy, dz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < pges; ++p)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                idx_t ind = at(r, c, p, rows, cols, pges);
                for (int a = 0; a < 3; ++a) {
                    double v[7], df, db;
                    line7(lsf, r, c, p, a, rows, cols, pges, v);
                    or_wenork3_derivative_boundary(&df, &db, v, R[a][ind], L[a][ind], sp[a], 0.0, 0.0, 1);
                    out[a][ind] = (fabs(v[4]) < fabs(v[2])) ? df : db;
                }
            }
}

/* WENORK3Reinitialization: +SD/@SDF3/WENORK3Reinitialization.m:3-61.  mask =
 * F<0, cubic distances of the argument, deltat = 0.3*min, RK3 (:30-57) */
void or_wenork3_reinit(double *F, int iterations, int rows, int cols, int pges, double dx, double dy,
                       double dz)
{
    idx_t n = (idx_t)rows * cols * pges;
    uint8_t *mask = (uint8_t *)malloc((size_t)n);
    double *d6 = dalloc(6 * n), *md = dalloc(n), *dt = dalloc(n), *step = dalloc(n), *f1 = dalloc(n),
           *f2 = dalloc(n);
    double *xpr = d6, *xpl = d6 + n, *ypf = d6 + 2 * n, *ypb = d6 + 3 * n, *zpu = d6 + 4 * n, *zpd = d6 + 5 * n;
    for (idx_t i = 0; i < n; ++i) mask[i] = F[i] < 0;
    or_set_distance(3, F, xpr, xpl, ypf, ypb, zpu, zpd, md, rows, cols, pges, dx, dy, dz);
    for (idx_t i = 0; i < n; ++i) dt[i] = 0.3 * md[i];
#define RS(src) or_wenork3_re_step(step, src, mask, dt, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, \
                                   dy, dz, (int)n)
    for (int it = 0; it < iterations; ++it) {
        RS(F);
        sub_into(f1, F, step, n);
        RS(f1);
        rk3_combine_stage2(f2, F, f1, step, n);
        RS(f2);
        rk3_combine_stage3(F, f2, step, n);
    }
#undef RS
    free(mask); free(d6); free(md); free(dt); free(step); free(f1); free(f2);
}

/* ENORK2Extend (+SD/@SDF3/ENORK2Extend.m:9-81, weno == 0: quadratic-ENO
 * distances, flat gradients floored at 1e-6, Heun RK2 :68-79) and
 * WENORK3Extend (+SD/@SDF3/WENORK3Extend.m:9-98, weno == 1: cubic distances,
 * max(|grad|,1e-6), RK3 :69-96).  C is extended in place off the zero set of F. */
static void variant_extend(int weno, double *C, int iterations, const double *F, int rows, int cols, int pges,
                           double dx, double dy, double dz)
{
    idx_t n = (idx_t)rows * cols * pges;
    double *d6 = dalloc(6 * n), *md = dalloc(n), *dt = dalloc(n), *v = dalloc(3 * n), *cp = dalloc(6 * n),
           *step = dalloc(n), *c1 = dalloc(n), *c2 = dalloc(n);
    double *xpr = d6, *xpl = d6 + n, *ypf = d6 + 2 * n, *ypb = d6 + 3 * n, *zpu = d6 + 4 * n, *zpd = d6 + 5 * n;
    double *vx = v, *vy = v + n, *vz = v + 2 * n;
    or_set_distance(weno ? 3 : 2, F, xpr, xpl, ypf, ypb, zpu, zpd, md, rows, cols, pges, dx, dy, dz);
    if (weno) or_wenork3_upwind_normal(vx, vy, vz, F, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, (int)n);
    else or_enork2_upwind_normal(vx, vy, vz, F, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, (int)n);
    for (idx_t i = 0; i < n; ++i) {
        double g = dmax2(sqrt(sq(vx[i]) + sq(vy[i]) + sq(vz[i])), 1e-6);
        double s = msign(F[i]);
        vx[i] = s * (vx[i] / g);
        vy[i] = s * (vy[i] / g);
        vz[i] = s * (vz[i] / g);
    }
    variant_interpolate(weno, cp, cp + n, cp + 2 * n, cp + 3 * n, cp + 4 * n, cp + 5 * n, xpr, xpl, ypf, ypb,
                        zpu, zpd, C, rows, cols, pges, dx, dy, dz);
    for (idx_t i = 0; i < n; ++i) dt[i] = 0.3 * md[i];
#define EXT(src) variant_extend_step(weno, step, dt, src, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cp, cp + n, \
                                     cp + 2 * n, cp + 3 * n, cp + 4 * n, cp + 5 * n, rows, cols, pges, dx, dy, dz)
    for (int it = 0; it < iterations; ++it) {
        if (weno) {
            EXT(C);
            sub_into(c1, C, step, n);
            EXT(c1);
            rk3_combine_stage2(c2, C, c1, step, n);
            EXT(c2);
            rk3_combine_stage3(C, c2, step, n);
        } else {
            EXT(C);
            sub_into(c1, C, step, n);
            EXT(c1);
            for (idx_t i = 0; i < n; ++i) C[i] = (C[i] + c1[i] - step[i]) / 2;
        }
    }
#undef EXT
    free(d6); free(md); free(dt); free(v); free(cp); free(step); free(c1); free(c2);
}

void or_enork2_extend(double *C, int iterations, const double *F, int rows, int cols, int pges, double dx,
                      double dy, double dz)
{
    variant_extend(0, C, iterations, F, rows, cols, pges, dx, dy, dz);
}

void or_wenork3_extend(double *C, int iterations, const double *F, int rows, int cols, int pges, double dx,
                       double dy, double dz)
{
    variant_extend(1, C, iterations, F, rows, cols, pges, dx, dy, dz);
}
