"""Planted mistakes in the oracle, each caught by a check that reads no
recorded output: the evidence that the closed forms, known answers and
invariants of test_oracle_math.py, test_oracle_kats.py and
test_oracle_relax.py pin every function of the oracle.

For each mutant a copy of the oracle package is made with one edit (a dropped
term, a wrong sign, index, side or slot, a wrong constant), the C library is
rebuilt from the edited source, and the named checks are run against the copy
in a child pytest (conftest.py puts VLS_ORACLE_ROOT first on sys.path).  The
mutant must make at least one of them fail.  All mutants are built and run
in parallel, once, in a module fixture.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
C_SRC = "ls_oracle.c"

MATH = "tests/test_oracle_math.py"
KATS = "tests/test_oracle_kats.py"
RELAX = "tests/test_oracle_relax.py"

# (id, file, original text, replacement, checks that must fail)
MUTANTS = [
    # Lagrange -> monomial coefficients and evaluation (shared_utilities.hpp:37-103)
    ("poly5_interp_sign", C_SRC, "t[k] = c[i][k] - zeros[i][j] * c[i][k - 1];",
     "t[k] = c[i][k] + zeros[i][j] * c[i][k - 1];",
     [f"{MATH}::test_six_distance_finds_the_root_of_a_quintic", f"{KATS}::test_boundary_interpolant_is_exact_for_quintics"]),
    ("poly5_eval_power", C_SRC, "        p *= x;\n", "        p *= x + 1e-3;\n",
     [f"{MATH}::test_six_distance_finds_the_root_of_a_quintic"]),
    # Brent root (shared_utilities.hpp:109-173), six_distance (boundary_location.cu:53-80)
    ("zeroin_tolerance", C_SRC, "zeroin_poly5(coef, 2 * s, 3 * s, 3e-16)", "zeroin_poly5(coef, 2 * s, 3 * s, 3e-6)",
     [f"{MATH}::test_six_distance_finds_the_root_of_a_quintic"]),
    ("zeroin_bracket", C_SRC, "            c = a; fc = fa;\n            d = b - a; e = d;",
     "            c = b; fc = fb;\n            d = b - a; e = d;",
     [f"{MATH}::test_six_distance_finds_the_root_of_a_quintic", f"{KATS}::test_boundary_location_exact_for_polynomials"]),
    ("six_distance_offset", C_SRC, "    return xc - 2 * s;\n", "    return xc - 1.9 * s;\n",
     [f"{MATH}::test_six_distance_finds_the_root_of_a_quintic"]),
    ("six_distance_kink_rule", C_SRC, "if (v1 * v2 < 0 || v3 * v4 < 0) return or_eno_quadratic_distance(v1, v2, v3, v4, s);",
     "if (v3 * v4 < 0) return or_eno_quadratic_distance(v1, v2, v3, v4, s);",
     [f"{MATH}::test_kink_falls_back_to_quadratic_eno"]),
    # quadratic ENO and cubic crossings (boundary_location.cu:6-20, weno_rk3_reinitialization.cu:74-117)
    ("eno_quadratic_root_side", C_SRC, "ksign(v0 - v2) * sqrt(disc)", "ksign(v2 - v0) * sqrt(disc)",
     [f"{MATH}::test_quadratic_eno_distance_exact_for_quadratics"]),
    ("cubic_coefficient", C_SRC, "    double c3 = (1.0 / 2.0 * (v1 - v0) - 1.0 / 2.0 * (v2 - v0) + 1.0 / 6.0 * (v3 - v0)) / (s * s * s);",
     "    double c3 = (1.0 / 2.0 * (v1 - v0) - 1.0 / 2.0 * (v2 - v0) + 1.0 / 3.0 * (v3 - v0)) / (s * s * s);",
     [f"{MATH}::test_cubic_distance_finds_the_root_of_a_cubic"]),
    ("cubic_kink_rule", C_SRC, "if (v0 * v1 < 0 || v2 * v3 < 0) return or_eno_quadratic_distance(v0, v1, v2, v3, s);",
     "if (v2 * v3 < 0) return or_eno_quadratic_distance(v0, v1, v2, v3, s);",
     [f"{MATH}::test_cubic_distance_kink_falls_back_to_quadratic_eno"]),
    # the distance arrays (boundary_location.cu:84-132)
    ("l_side_distance", C_SRC, "left_out[axis][nb] = ds - d;", "left_out[axis][nb] = d;",
     [f"{KATS}::test_boundary_location_slots_and_roots"]),
    ("cubic_order_uses_eno", C_SRC, ": (order == 3) ? or_cubic_distance(v[2], f0, f2, v[5], ds)",
     ": (order == 3) ? or_eno_quadratic_distance(v[2], f0, f2, v[5], ds)",
     [f"{KATS}::test_boundary_location_exact_for_polynomials"]),
    ("crossing_axis", C_SRC, "if (axis == 0) cc += 1; else if (axis == 1) rr += 1; else pp += 1;",
     "if (axis == 0) rr += 1; else if (axis == 1) cc += 1; else pp += 1;",
     [f"{KATS}::test_boundary_location_slots_and_roots"]),
    # uniform WENO5 (re_step.cu:11-37): stencil-weight forms and the Jiang-Peng pair
    ("weno_linear_weights", C_SRC, "w1 = 0.1 * (a2 * a3), w2 = 0.6 * (a1 * a3)", "w1 = 0.2 * (a2 * a3), w2 = 0.5 * (a1 * a3)",
     [f"{MATH}::test_weno5_pair_matches_both_stencil_weight_forms"]),
    ("weno_reference_eps", C_SRC, "double a1 = 0.1 / sq(S1 + eps);", "double a1 = 0.1 / sq(S1 + 10 * eps);",
     [f"{MATH}::test_weno5_pair_matches_both_stencil_weight_forms", f"{MATH}::test_weno5_single_division_form_matches_reference_order"]),
    ("jp_weight", C_SRC, "w1 = 6.0 * (a0 * a2), w2 = 3.0 * (a0 * a1);", "w1 = 5.0 * (a0 * a2), w2 = 3.0 * (a0 * a1);",
     [f"{KATS}::test_weno5_is_fifth_order_on_smooth_data", f"{MATH}::test_weno5_pair_matches_both_stencil_weight_forms"]),
    ("jp_eps_scale", C_SRC, "e = 12.0 / 13.0 * 1e-6 * ds * ds;", "e = 1e-6 * ds * ds;",
     [f"{MATH}::test_weno5_pair_matches_both_stencil_weight_forms"]),
    ("jp_fore_sign", C_SRC, "*d_fore = (c + jp_q(D[4], D[3], D[2], D[1], e)) * k;",
     "*d_fore = (c - jp_q(D[4], D[3], D[2], D[1], e)) * k;",
     [f"{MATH}::test_weno5_pair_matches_both_stencil_weight_forms", f"{KATS}::test_upwind_normal_takes_the_side_with_the_smaller_neighbour"]),
    # nonuniform WENO (re_step.cu:41-243)
    ("nonuniform_linear_weight", C_SRC,
     "        c0 = ((xm05 - xm15) / (x25 - xm25)) * ((xm05 - xm25) / (x25 - xm15));",
     "        c0 = ((xm05 - xm15) / (x25 - xm25)) * ((xm05 - xm15) / (x25 - xm15));",
     [f"{KATS}::test_weno5_is_fifth_order_on_smooth_data", f"{MATH}::test_nonuniform_weno_on_uniform_stencil_is_weno5"]),
    ("nonuniform_candidate", C_SRC,
     "        v1 = u0 + (um1 - u0) * ((x05 - xm05) / (x15 - xm15)) * ((x15 - xm05) / (x05 - xm15))",
     "        v1 = u0 + (um1 - u0) * ((x05 - xm05) / (x15 - xm15)) * ((x15 - xm15) / (x05 - xm15))",
     [f"{MATH}::test_nonuniform_weno_exact_for_cubics"]),
    ("nonuniform_smoothness", C_SRC,
     "    double s1 = 4.0 * sq((x05 - xm05) / (x15 - xm15)) *",
     "    double s1 = 40.0 * sq((x05 - xm05) / (x15 - xm15)) *",
     [f"{KATS}::test_weno5_is_fifth_order_on_smooth_data", f"{MATH}::test_nonuniform_weno_on_uniform_stencil_is_weno5"]),
    ("nonuniform_stencil_node", C_SRC, "        x[4] = r1; x[5] = ds; x[6] = 2 * ds;\n        h[4] = h_fore; h[5] = p[4]; h[6] = p[5];\n    } else {\n        x[4] = ds; x[5] = 2 * ds; x[6] = 3 * ds;\n        h[4] = p[4]; h[5] = p[5]; h[6] = p[6];\n    }\n    if (l1 < ds) {\n        x[2] = -l1; x[1] = -ds; x[0] = -2 * ds;\n        h[2] = h_back; h[1] = p[2]; h[0] = p[1];\n    } else {\n        x[2] = -ds; x[1] = -2 * ds; x[0] = -3 * ds;\n        h[2] = p[2]; h[1] = p[1]; h[0] = p[0];\n    }\n    /* first",
     "        x[4] = r1; x[5] = ds; x[6] = 2 * ds;\n        h[4] = h_fore; h[5] = p[4]; h[6] = p[5];\n    } else {\n        x[4] = ds; x[5] = 2 * ds; x[6] = 3 * ds;\n        h[4] = p[4]; h[5] = p[5]; h[6] = p[6];\n    }\n    if (l1 < ds) {\n        x[2] = -l1; x[1] = -ds; x[0] = -2 * ds;\n        h[2] = h_back; h[1] = p[1]; h[0] = p[0];\n    } else {\n        x[2] = -ds; x[1] = -2 * ds; x[0] = -3 * ds;\n        h[2] = p[2]; h[1] = p[1]; h[0] = p[0];\n    }\n    /* first",
     [f"{MATH}::test_nonuniform_weno_exact_for_cubics"]),
    ("reinit_interface_test", C_SRC, "int cross = rule == 0 ? (p[2] * p[3] < 0 || p[3] * p[4] < 0) : (r1 < ds || l1 < ds);\n    if (cross) {\n        if (rule == 0) { v_fore = 0.0; v_back = 0.0; }\n        weno_nonuniform_pair",
     "int cross = rule == 0 ? (p[3] * p[4] < 0) : (r1 < ds || l1 < ds);\n    if (cross) {\n        if (rule == 0) { v_fore = 0.0; v_back = 0.0; }\n        weno_nonuniform_pair",
     [f"{KATS}::test_re_step_interface_branch_follows_the_sign_test"]),
    # Godunov step and re_step (re_step.cu:283-393)
    ("godunov_inside_branch", C_SRC, "        return (sqrt(dmax2(sq(dmin2(0.0, xL)), sq(dmax2(0.0, xR)))",
     "        return (sqrt(dmax2(sq(dmax2(0.0, xL)), sq(dmax2(0.0, xR)))",
     [f"{KATS}::test_re_step_exact_on_planes_including_interface_nodes"]),
    ("frozen_threshold", C_SRC, "void or_re_step(double *step, const double *lsf, const uint8_t *mask, const double *deltat,\n                SIX_DIST_ARGS, int rows, int cols, int pges, double dx, double dy, double dz,\n                int num_ele)\n{\n    (void)num_ele;\n    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};\n    const double sp[3] = {dx, dy, dz};\n#pragma omp parallel for collapse(2) schedule(static)\n    for (int p = 0; p < pges; ++p)\n        for (int c = 0; c < cols; ++c)\n            for (int r = 0; r < rows; ++r) {\n                idx_t ind = at(r, c, p, rows, cols, pges);\n                double eps = 1e-6 * dx;",
     "void or_re_step(double *step, const double *lsf, const uint8_t *mask, const double *deltat,\n                SIX_DIST_ARGS, int rows, int cols, int pges, double dx, double dy, double dz,\n                int num_ele)\n{\n    (void)num_ele;\n    const double *R[3] = {xpr, ypf, zpu}, *L[3] = {xpl, ypb, zpd};\n    const double sp[3] = {dx, dy, dz};\n#pragma omp parallel for collapse(2) schedule(static)\n    for (int p = 0; p < pges; ++p)\n        for (int c = 0; c < cols; ++c)\n            for (int r = 0; r < rows; ++r) {\n                idx_t ind = at(r, c, p, rows, cols, pges);\n                double eps = 1e-6 * dz;",
     [f"{KATS}::test_frozen_points"]),
    # ENO2 (boundary_correction.cu:80-200)
    ("eno2_subcell_value", C_SRC, "    double vr = (pr == ds) ? v2 : 0;\n    *sR", "    double vr = v2;\n    *sR",
     [f"{KATS}::test_re_step_exact_on_planes_including_interface_nodes"]),
    ("eno2_minmod", C_SRC, "    p2m = 0.5 * min_mod(p2, p2l) / sq(ds);", "    p2m = 0.5 * min_mod(p2, p2) / sq(ds);",
     [f"{KATS}::test_variant_upwind_normals_take_the_same_side", f"{MATH}::test_eno2_reinit_reduces_gradient_error"]),
    # upwind_normal (upwind_normal.cu:4-104)
    ("upwind_normal_side", C_SRC, "                    or_weno_derivative_boundary(&df, &db, v, R[a][ind], L[a][ind], sp[a], 0.0,\n                                                0.0, 1);\n                    out[a][ind] = (fabs(v[4]) < fabs(v[2])) ? df : db;",
     "                    or_weno_derivative_boundary(&df, &db, v, R[a][ind], L[a][ind], sp[a], 0.0,\n                                                0.0, 1);\n                    out[a][ind] = (fabs(v[4]) < fabs(v[2])) ? db : df;",
     [f"{KATS}::test_upwind_normal_takes_the_side_with_the_smaller_neighbour"]),
    # boundary_interpolant (boundary_interpolant.cu:77-141)
    ("interpolant_point", C_SRC, "*c_fore = poly5_eval(coef, 3 * s + dis_f);", "*c_fore = poly5_eval(coef, 3 * s - dis_f);",
     [f"{KATS}::test_boundary_interpolant_is_exact_for_quintics"]),
    ("interpolant_slots", C_SRC, "sixth_interp(&CF[a][ind], &CB[a][ind], R[a][ind], L[a][ind], sp[a], v);",
     "sixth_interp(&CB[a][ind], &CF[a][ind], R[a][ind], L[a][ind], sp[a], v);",
     [f"{KATS}::test_boundary_interpolant_is_exact_for_quintics"]),
    # extend_step (extend_step.cu:10-125)
    ("extend_upwind", C_SRC, "                    terms[2 * a] = dmin2(0.0, vel) * df;\n                    terms[2 * a + 1] = dmax2(0.0, vel) * db;",
     "                    terms[2 * a] = dmin2(0.0, vel) * db;\n                    terms[2 * a + 1] = dmax2(0.0, vel) * df;",
     [f"{KATS}::test_advection_and_extension_steps_are_upwind"]),
    ("extend_interface_values", C_SRC, "                                                CF[a][ind], CB[a][ind], 1);\n                    double vel = V[a][ind];",
     "                                                CB[a][ind], CF[a][ind], 1);\n                    double vel = V[a][ind];",
     [f"{KATS}::test_extend_step_at_the_interface"]),
    # advection_step, weno_derivative (advection_step.cu:30-56, nonoscillating_interpolant.cu:59-122)
    ("advection_upwind", C_SRC, "                    t[2 * a] = dmin2(0.0, V[a][ind]) * df;", "                    t[2 * a] = dmax2(0.0, V[a][ind]) * df;",
     [f"{KATS}::test_advection_and_extension_steps_are_upwind"]),
    ("weno_derivative_slots", C_SRC, "plain_pair(&F[a][ind], &B[a][ind], lsf", "plain_pair(&B[a][ind], &F[a][ind], lsf",
     [f"{KATS}::test_weno_derivative_slots"]),
    # set_calculus_toolbox (set_calculus_toolbox.cu:9-107)
    ("toolbox_cross_denominator", C_SRC, "double fxy = (F_(1, 1, 0) + F_(-1, -1, 0) - F_(1, -1, 0) - F_(-1, 1, 0)) / (4 * ds * ds);",
     "double fxy = (F_(1, 1, 0) + F_(-1, -1, 0) - F_(1, -1, 0) - F_(-1, 1, 0)) / (4 * dx * dy);",
     [f"{KATS}::test_set_calculus_toolbox_on_a_quadratic"]),
    ("toolbox_mean_curvature", C_SRC, "out[14][ind] = -lap / g + (fx * c1 + fy * c2 + fz * c3) / (g * g * g);",
     "out[14][ind] = lap / g - (fx * c1 + fy * c2 + fz * c3) / (g * g * g);",
     [f"{KATS}::test_set_calculus_toolbox_on_a_quadratic"]),
    ("toolbox_gaussian_curvature", C_SRC, "c2 = (fyz * fzx - fxy * fzz) * fx + (fxx * fzz - fzx * fzx) * fy",
     "c2 = (fyz * fzx - fxy * fzz) * fx + (fxx * fzz - fxy * fxy) * fy",
     [f"{KATS}::test_set_calculus_toolbox_on_a_quadratic"]),
    ("toolbox_heaviside", C_SRC, "out[16][ind] = dot / (g * g);", "out[16][ind] = dot / g;",
     [f"{KATS}::test_set_calculus_toolbox_on_a_quadratic"]),
    # auxi_set_calculus_toolbox (auxi_set_calculus_toolbox.cu:9-199) and its MATLAB tail
    ("aux_cross_order", C_SRC, "cross3(C, fx, fy, fz, ax, ay, az);", "cross3(C, ax, ay, az, fx, fy, fz);",
     [f"{MATH}::test_aux_toolbox_frame_and_curvatures_match_the_analytic_curve"]),
    ("aux_geodesic_term", C_SRC, "out[21][ind] = (n[0] * (w1[0] + w2[0]) + n[1] * (w1[1] + w2[1]) + n[2] * (w1[2] + w2[2])) / nc;",
     "out[21][ind] = (n[0] * w1[0] + n[1] * w1[1] + n[2] * w1[2]) / nc;",
     [f"{MATH}::test_aux_toolbox_frame_and_curvatures_match_the_analytic_curve"]),
    ("aux_heaviside", C_SRC, "out[25][ind] = dot / (ag * ag);", "out[25][ind] = dot / ag;",
     [f"{MATH}::test_aux_heaviside_and_delta_away_from_the_aux_surface"]),
    ("aux_tail_wrap_axis", C_SRC, "idx_t yp = wrap_at(r + 1, c, p, rows, cols, pges), ym = wrap_at(r - 1, c, p, rows, cols, pges);",
     "idx_t yp = wrap_at(r, c + 1, p, rows, cols, pges), ym = wrap_at(r, c - 1, p, rows, cols, pges);",
     [f"{MATH}::test_aux_toolbox_tail_is_gd3_periodic_differencing"]),
    ("aux_tail_cross_order", C_SRC, "cross3(ch, h2[0], h2[1], h2[2], h1[0], h1[1], h1[2]);",
     "cross3(ch, h1[0], h1[1], h1[2], h2[0], h2[1], h2[2]);",
     [f"{MATH}::test_aux_toolbox_tail_is_gd3_periodic_differencing", f"{MATH}::test_aux_dirac_delta_integrates_to_the_curve_length"]),
    # reductions (SDF3.m:73-85, src/0_0_setup/+SD/@SDF3/SDF3.m:126-136)
    ("neumaier_compensation", C_SRC, "    if (fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x;", "    if (fabs(k->s) >= fabs(x)) k->c += 0.0;",
     [f"{MATH}::test_compensated_sums_against_fsum"]),
    ("towers_area_mask", C_SRC, "if (fabs(F[i]) < 2 * ds) kadd(&sa, out[17][i]);", "if (fabs(F[i]) < ds) kadd(&sa, out[17][i]);",
     [f"{MATH}::test_towers_integrals_of_a_scaled_distance"]),
    ("cosine_delta_half", C_SRC, "delta = g * (1 + cos(pi * v / eps)) / eps / 2;", "delta = g * (1 + cos(pi * v / eps)) / eps;",
     [f"{MATH}::test_cosine_area_volume_match_closed_forms"]),
    ("cosine_heaviside", C_SRC, "heav = (1 + v / eps + sin(pi * v / eps) / pi) / 2;", "heav = (1 + v / eps - sin(pi * v / eps) / pi) / 2;",
     [f"{MATH}::test_cosine_area_volume_match_closed_forms"]),
    # MATLAB wrappers: setDistance, the RK loops, extension velocity, advection
    ("min_dist", C_SRC, "min_dist[i] = dmin2(dmin2(xpr[i], xpl[i]), dmin2(dmin2(ypf[i], ypb[i]), dmin2(zpu[i], zpd[i])));",
     "min_dist[i] = dmin2(dmin2(xpr[i], xpl[i]), dmin2(dmin2(ypf[i], ypb[i]), zpu[i]));",
     [f"{KATS}::test_set_distance_min_dist"]),
    ("rk3_stage2", C_SRC, "out[i] = 3. / 4. * fn[i] + 1. / 4. * t2;", "out[i] = 2. / 3. * fn[i] + 1. / 3. * t2;",
     [f"{MATH}::test_rk3_is_third_order_in_time"]),
    ("rk3_stage3", C_SRC, "fn[i] = 1. / 3. * fn[i] + 2. / 3. * t15;", "fn[i] = 1. / 4. * fn[i] + 3. / 4. * t15;",
     [f"{MATH}::test_rk3_is_third_order_in_time"]),
    ("reinit_time_step", C_SRC, "for (idx_t i = 0; i < n; ++i) { mask[i] = F[i] < 0; dt[i] = 0.3 * min_dist[i]; }",
     "for (idx_t i = 0; i < n; ++i) { mask[i] = F[i] < 0; dt[i] = 0.5 * min_dist[i]; }",
     [f"{KATS}::test_one_reinit_iteration_far_from_the_interface"]),
    ("eno2_heun", C_SRC, "for (idx_t i = 0; i < n; ++i) F[i] = (F[i] + ft[i] - step[i]) / 2;",
     "for (idx_t i = 0; i < n; ++i) F[i] = (F[i] + ft[i] + step[i]) / 2;",
     [f"{KATS}::test_one_reinit_iteration_far_from_the_interface"]),
    ("extension_velocity_sign", C_SRC, "        double s = msign(F[i]);\n        vx[i] = s * (vx[i] / g);\n        vy[i] = s * (vy[i] / g);\n        vz[i] = s * (vz[i] / g);\n    }\n}",
     "        double s = ksign(F[i]);\n        vx[i] = s * (vx[i] / g);\n        vy[i] = s * (vy[i] / g);\n        vz[i] = s * (vz[i] / g);\n    }\n}",
     [f"{KATS}::test_extension_velocity_is_zero_on_exact_zeros"]),
    ("extension_floor", C_SRC, "or_extension_velocity(F, xpr, xpl, ypf, ypb, zpu, zpd, 1e-6, v,", "or_extension_velocity(F, xpr, xpl, ypf, ypb, zpu, zpd, 1e-14, v,",
     [f"{KATS}::test_one_extension_iteration_far_from_the_interface"]),
    ("extension_time_step", C_SRC, "    for (idx_t i = 0; i < n; ++i) dt[i] = 0.3 * min_dist[i];\n    or_extend_loop",
     "    for (idx_t i = 0; i < n; ++i) dt[i] = 0.25 * min_dist[i];\n    or_extend_loop",
     [f"{KATS}::test_one_extension_iteration_far_from_the_interface"]),
    ("redistance_distances", C_SRC, "#define RS(src) or_re_step(step, src, mask, adt, D6(ad),", "#define RS(src) or_re_step(step, src, mask, adt, D6(fd),",
     [f"{MATH}::test_surface_redistance_keeps_the_zero_set"]),
    ("advect_dt", C_SRC, "        or_advection_step(a, vx, vy, vz, f1, rows, cols, pges, dx, dy, dz);\n        for (idx_t i = 0; i < n; ++i) step[i] = dt * a[i];",
     "        or_advection_step(a, vx, vy, vz, f1, rows, cols, pges, dx, dy, dz);\n        for (idx_t i = 0; i < n; ++i) step[i] = 0.5 * dt * a[i];",
     [f"{MATH}::test_rk3_is_third_order_in_time", f"{MATH}::test_advection_translates_a_sphere"]),
    ("normal_velocity_axis", C_SRC, "out[ind] = vx[ind] * (fx / g) + vy[ind] * (fy / g) + vz[ind] * (fz / g);",
     "out[ind] = vx[ind] * (fy / g) + vy[ind] * (fx / g) + vz[ind] * (fz / g);",
     [f"{KATS}::test_normal_velocity_of_a_linear_field"]),
    # WENO-RK3 / ENO-RK2 variants
    ("eno_cubic_correction", C_SRC, "double a1 = (x[3] - x[2]) * (x[3] - x[1]) * min_abs(d3_m05, d3_m15);",
     "double a1 = (x[3] - x[2]) * (x[3] - x[0]) * min_abs(d3_m05, d3_m15);",
     [f"{MATH}::test_eno_cubic_derivative_exact_for_cubics"]),
    ("wenork3_interface_test", C_SRC, "int cross = rule == 0 ? (p[2] * p[3] < 0 || p[3] * p[4] < 0) : (r1 < ds || l1 < ds);\n    if (cross) {\n        double x[7], h[7];",
     "int cross = rule == 0 ? (p[3] * p[4] < 0) : (r1 < ds || l1 < ds);\n    if (cross) {\n        double x[7], h[7];",
     [f"{MATH}::test_wenork3_derivative_boundary_branches", f"{KATS}::test_re_step_interface_branch_follows_the_sign_test"]),
    ("wenork3_normal_side", C_SRC, "                    or_wenork3_derivative_boundary(&df, &db, v, R[a][ind], L[a][ind], sp[a], 0.0, 0.0, 1);\n                    out[a][ind] = (fabs(v[4]) < fabs(v[2])) ? df : db;",
     "                    or_wenork3_derivative_boundary(&df, &db, v, R[a][ind], L[a][ind], sp[a], 0.0, 0.0, 1);\n                    out[a][ind] = (fabs(v[4]) <= fabs(v[2])) ? db : df;",
     [f"{KATS}::test_variant_upwind_normals_take_the_same_side"]),
    ("enork2_normal_side", C_SRC, "out[a][ind] = (fabs(vr) < fabs(vl)) ? sR : sL;", "out[a][ind] = (fabs(vr) < fabs(vl)) ? sL : sR;",
     [f"{KATS}::test_variant_upwind_normals_take_the_same_side"]),
    ("eno2_field_interface_value", C_SRC, "    double v_f = (dis_f == ds) ? v2 : v_fore;", "    double v_f = v2;",
     [f"{KATS}::test_variant_extend_step_at_the_interface"]),
    ("eno2_interface_values", C_SRC, "        double x = dis_f - ds / 2;", "        double x = dis_f + ds / 2;",
     [f"{MATH}::test_variant_interface_values_exact_for_polynomials"]),
    ("cubic_interface_values", C_SRC, "        double xc = 2 * ds - dis_b;", "        double xc = 2 * ds + dis_b;",
     [f"{MATH}::test_variant_interface_values_exact_for_polynomials"]),
    ("variant_upwind_sum", C_SRC, "return (dmin2(0.0, vx) * fo[0] + dmax2(0.0, vx) * bk[0] + dmin2(0.0, vy) * fo[1] +",
     "return (dmin2(0.0, vx) * bk[0] + dmax2(0.0, vx) * fo[0] + dmin2(0.0, vy) * fo[1] +",
     [f"{KATS}::test_advection_and_extension_steps_are_upwind"]),
    ("wenork3_frozen", C_SRC, "if (weno && (xpr[ind] < eps ||", "if (!weno && (xpr[ind] < eps ||",
     [f"{KATS}::test_frozen_points"]),
    ("wenork3_reinit_order", C_SRC, "    or_set_distance(3, F, xpr, xpl, ypf, ypb, zpu, zpd, md, rows, cols, pges, dx, dy, dz);\n    for (idx_t i = 0; i < n; ++i) dt[i] = 0.3 * md[i];\n#define RS",
     "    or_set_distance(3, F, xpr, xpl, ypf, ypb, zpu, zpd, md, rows, cols, pges, dx, dy, dz);\n    for (idx_t i = 0; i < n; ++i) dt[i] = 0.4 * md[i];\n#define RS",
     [f"{KATS}::test_one_reinit_iteration_far_from_the_interface"]),
    ("variant_extend_heun", C_SRC, "for (idx_t i = 0; i < n; ++i) C[i] = (C[i] + c1[i] - step[i]) / 2;",
     "for (idx_t i = 0; i < n; ++i) C[i] = (C[i] + c1[i]) / 2;",
     [f"{KATS}::test_one_extension_iteration_far_from_the_interface"]),
    ("variant_extend_floor", C_SRC, "        double g = dmax2(sqrt(sq(vx[i]) + sq(vy[i]) + sq(vz[i])), 1e-6);\n        double s = msign(F[i]);\n        vx[i] = s * (vx[i] / g);\n        vy[i] = s * (vy[i] / g);\n        vz[i] = s * (vz[i] / g);\n    }\n    variant_interpolate",
     "        double g = dmax2(sqrt(sq(vx[i]) + sq(vy[i]) + sq(vz[i])), 1e-14);\n        double s = msign(F[i]);\n        vx[i] = s * (vx[i] / g);\n        vy[i] = s * (vy[i] / g);\n        vz[i] = s * (vz[i] / g);\n    }\n    variant_interpolate",
     [f"{KATS}::test_one_extension_iteration_far_from_the_interface"]),
    # the relaxation step's smoothers and tension/pressure (oracle/relax_oracle.py)
    ("smooth_fft_symbol", "relax_oracle.py", "    sym = k2 if diffusion else k2 ** 2\n", "    sym = k2 if diffusion else k2 ** 1.5\n",
     [f"{RELAX}::test_smooth_fft_scales_a_fourier_mode"]),
    ("fd_eigenvalues", "relax_oracle.py", "return -4 * np.sin(math.pi * np.arange(n) / n) ** 2 / h ** 2",
     "return -4 * np.sin(math.pi * np.arange(n) / n) ** 2 / h",
     [f"{RELAX}::test_exact_solve_and_gmres_solve_the_sparse_system"]),
]


def _build(tmp, mid, fname, old, new):
    """the oracle package copied to tmp/<mid>/oracle with one edit; returns the
    directory to put first on sys.path"""
    root = os.path.join(tmp, mid)
    pkg = os.path.join(root, "oracle")
    os.makedirs(pkg)
    for f in os.listdir(ORACLE):
        if f.endswith((".py", ".c")):
            shutil.copy(os.path.join(ORACLE, f), pkg)
    path = os.path.join(pkg, fname)
    text = open(path).read()
    n = text.count(old)
    if n != 1:
        raise AssertionError(f"{mid}: the original text occurs {n} times in {fname}")
    open(path, "w").write(text.replace(old, new))
    if fname == C_SRC:
        subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fPIC", "-fopenmp", "-shared", "-o",
                        os.path.join(pkg, "liblsoracle.so"), path, "-lm"], check=True, capture_output=True)
    else:
        shutil.copy(os.path.join(ORACLE, "liblsoracle.so"), pkg)
    return root


def _run(tmp, mutant):
    mid, fname, old, new, checks = mutant
    try:
        root = _build(tmp, mid, fname, old, new)
    except (AssertionError, subprocess.CalledProcessError) as e:
        return -1, f"could not plant the mistake: {e}"
    env = dict(os.environ, VLS_ORACLE_ROOT=root, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + checks,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-1500:]


@pytest.fixture(scope="module")
def outcomes():
    with tempfile.TemporaryDirectory() as tmp:
        workers = max(1, min(8, os.cpu_count() or 1))
        with concurrent.futures.ThreadPoolExecutor(workers) as ex:
            futs = {m[0]: ex.submit(_run, tmp, m) for m in MUTANTS}
            return {k: f.result() for k, f in futs.items()}


def test_unmutated_oracle_passes_the_checks():
    # the same child run against the unedited package: every named check passes
    checks = sorted({c for m in MUTANTS for c in m[4]})
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] + checks,
                       cwd=ROOT, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("mid", [m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(outcomes, mid):
    rc, out = outcomes[mid]
    # pytest exit code 1: some check failed (2-5 would be an error of the run itself)
    assert rc == 1, f"{mid} survived (exit {rc}):\n{out}"
