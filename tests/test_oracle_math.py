"""What the mathematics fixes about the oracle: closed forms, exactness on
polynomials, convergence to known solutions and invariants, for the kernels
and for the MATLAB wrappers (RK loops, extension, surface redistance,
integrals).  test_oracle_kats.py adds known answers on kinked, linear and
planar fields; test_oracle_mutants.py shows that a planted mistake in any
oracle function fails one of these checks.

Each check is chosen so that a plausible slip fails it: a wrong stencil
offset or WENO weight breaks polynomial exactness; a sign slip in the
Godunov flux or the mask breaks the plane and sphere tests; a wrong RK
coefficient breaks third-order convergence in time; a transposed axis breaks
the anisotropic ellipsoid checks; an off-by-one in the delta/H masks or the
weights breaks the area/volume limits.
"""
import math

import numpy as np
import pytest

import cases
from oracle import pyoracle as po

RNG = np.random.default_rng(20181207)


# --------------------------------------------------------------------------
# WENO5 derivatives (re_step.cu:11-37, 79-269)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("trial", range(5))
def test_weno5_exact_for_cubics(trial):
    a, b, c, d = RNG.normal(size=4)
    h = RNG.uniform(0.01, 0.2)
    x = (np.arange(7) - 3) * h
    p = a + b * x + c * x ** 2 + d * x ** 3
    v = np.diff(p) / h
    scale = abs(a) + abs(b) + abs(c) + abs(d)
    assert abs(po.weno_onesided(*v[:5]) - b) <= 1e-12 * scale / h * h + 1e-13 * scale
    assert abs(po.weno_onesided(v[5], v[4], v[3], v[2], v[1]) - b) <= 1e-13 * scale


def test_weno5_uses_the_upwind_side():
    # a kink at the node: each one-sided derivative sees only its own line
    h = 0.1
    x = (np.arange(7) - 3) * h
    p = np.where(x <= 0, 2.0 * x, -5.0 * x)
    df, db = po.weno_derivative_boundary(p, h, h, h, rule=1)
    assert abs(db - 2.0) < 1e-3 and abs(df + 5.0) < 1e-3


@pytest.mark.parametrize("trial", range(5))
def test_nonuniform_weno_exact_for_cubics(trial):
    a, b, c, d = RNG.normal(size=4)
    f = lambda t: a + b * t + c * t ** 2 + d * t ** 3  # noqa: E731
    h = 0.1
    x = (np.arange(7) - 3) * h
    r1, l1 = RNG.uniform(0.05, 0.95, size=2) * h
    df, db = po.weno_derivative_boundary(f(x), r1, l1, h, f(r1), f(-l1), rule=1)
    assert abs(df - b) < 1e-12 and abs(db - b) < 1e-12


def test_nonuniform_weno_on_uniform_stencil_is_weno5():
    h = 0.1
    x = (np.arange(7) - 3) * h
    p = np.sin(3 * x) + 0.3 * x + 0.05  # sign change between nodes 2 and 3
    assert p[2] * p[3] < 0
    df, db = po.weno_derivative_boundary(p, h, h, h, rule=0)
    v = np.diff(p) / h
    assert abs(df - po.weno_onesided(v[5], v[4], v[3], v[2], v[1])) < 1e-14
    assert abs(db - po.weno_onesided(*v[:5])) < 1e-14


def test_nonuniform_weno_at_full_spacing_is_weno5_across_kinks():
    """The reinitialization's interface path at a point with no subcell
    distance (re_step.cu:248-269 with xpr = xpl = dx) is the uniform WENO5 up
    to rounding, also where the stencil spans a kink of the distance function
    or a thin sheet: the FP32 bulk evaluates such points with the uniform form
    instead of deferring them."""
    rng = np.random.default_rng(11)
    h = 2.4 / 255
    x = (np.arange(7) - 3) * h
    worst = 0.0
    checked = 0
    for t in range(3000):
        a = rng.uniform(-1, 1) * h
        kind = t % 4
        if kind == 0:
            p = (x - a) * rng.uniform(0.2, 2)
        elif kind == 1:
            p = np.abs(x - a) - rng.uniform(0.05, 1.5) * h
        elif kind == 2:
            p = rng.uniform(0.05, 1.5) * h - np.abs(x - a)
        else:
            p = (x - a) + rng.uniform(-30, 30) * (x - a) ** 2 + rng.uniform(-300, 300) * (x - a) ** 3
        p = p + rng.normal(0, 1e-3 * h, size=7)
        if not (p[2] * p[3] < 0 or p[3] * p[4] < 0):
            continue
        checked += 1
        fu, bu = po.weno_pair(p, h)
        fn, bn = po.weno_derivative_boundary(p, h, h, h, rule=0)
        worst = max(worst, abs(fn - fu) / max(abs(fu), 1.0), abs(bn - bu) / max(abs(bu), 1.0))
    assert checked > 1000
    assert worst < 1e-14, worst


@pytest.mark.parametrize("trial", range(5))
def test_weno5_pair_exact_for_cubics(trial):
    a, b, c, d = RNG.normal(size=4)
    h = RNG.uniform(0.01, 0.2)
    x = (np.arange(7) - 3) * h
    p = a + b * x + c * x ** 2 + d * x ** 3
    scale = abs(a) + abs(b) + abs(c) + abs(d)
    df, db = po.weno_pair(p, h)
    assert abs(df - b) <= 1e-12 * scale and abs(db - b) <= 1e-12 * scale


def _pair_cases(n):
    """nodal lines at scales from smooth (indicators far below the 1e-6 of the
    weights) through eps-dominated to rough, with kinks and random data"""
    out = []
    for _ in range(n):
        h = RNG.uniform(0.005, 0.2)
        x = (np.arange(7) - 3) * h
        kind = RNG.integers(4)
        if kind == 0:  # linear plus noise of amplitude 10^k * h (k = -6..0)
            p = RNG.normal() + RNG.normal() * x + 10.0 ** RNG.integers(-6, 1) * h * RNG.uniform(-1, 1, 7)
        elif kind == 1:  # kink
            s = RNG.integers(0, 7)
            p = np.where(np.arange(7) < s, RNG.normal() * x, RNG.normal() * x) + RNG.normal()
        elif kind == 2:  # smooth
            p = np.sin(RNG.normal() * 3 * x + RNG.normal()) * 10.0 ** RNG.integers(-3, 3)
        else:  # random
            p = RNG.normal(size=7) * 10.0 ** RNG.integers(-3, 3)
        out.append((p, h))
    return out


def test_weno5_pair_matches_both_stencil_weight_forms():
    """the Jiang-Peng pair of the algorithms against the reference's own
    evaluation order and the single-division form, on the same divided
    differences: a dropped term, a wrong index or a wrong eps scaling shows"""
    import ctypes

    ref = po.lib().or_weno_onesided_reference
    ref.restype = ctypes.c_double
    for p, h in _pair_cases(400):
        v = np.diff(p) / h
        df, db = po.weno_pair(p, h)
        scale = max(1.0, np.abs(v).max())
        want_b = ref(*[ctypes.c_double(t) for t in v[:5]])
        want_f = ref(*[ctypes.c_double(t) for t in (v[5], v[4], v[3], v[2], v[1])])
        assert abs(db - want_b) <= 1e-13 * scale, (p, h, db, want_b)
        assert abs(df - want_f) <= 1e-13 * scale, (p, h, df, want_f)
        assert abs(db - po.weno_onesided(*v[:5])) <= 1e-13 * scale
        assert abs(df - po.weno_onesided(v[5], v[4], v[3], v[2], v[1])) <= 1e-13 * scale


def test_weno5_single_division_form_matches_reference_order():
    import ctypes

    ref = po.lib().or_weno_onesided_reference
    ref.restype = ctypes.c_double
    for _ in range(200):
        w = RNG.normal(size=5) * 10.0 ** RNG.integers(-3, 3)
        got = po.weno_onesided(*w)
        want = ref(*[ctypes.c_double(t) for t in w])
        assert abs(got - want) <= 1e-14 * max(1.0, np.abs(w).max())


# --------------------------------------------------------------------------
# subcell distances (boundary_location.cu:6-80, shared_utilities.hpp:37-173)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("trial", range(6))
def test_six_distance_finds_the_root_of_a_quintic(trial):
    s = 0.07
    root = RNG.uniform(2.05, 2.95) * s
    others = RNG.uniform(-3, -1, size=4) * s  # roots away from [0, 5s]
    xs = np.arange(6) * s
    y = (xs - root) * np.prod([xs - o for o in others], axis=0) * (1 + 0.1 * RNG.normal())
    dist = po.six_distance(*y, s)
    assert abs(dist - (root - 2 * s)) < 1e-13 * s


def test_six_distance_linear_data():
    s = 0.05
    v = 0.3 - 1.7 * np.arange(6) * s  # root at 0.3/1.7 ~ 0.176 in [2s, 3s]
    assert abs(po.six_distance(*v, s) - (0.3 / 1.7 - 2 * s)) < 1e-14


def test_quadratic_eno_distance_exact_for_quadratics():
    s = 0.1
    # f(t) = (t - 0.43 s)(t + 3.1 s) at t = -s, 0, s, 2s: root 0.43 s from node 0
    f = lambda t: (t - 0.43 * s) * (t + 3.1 * s)  # noqa: E731
    d = po.eno_quadratic_distance(f(-s), f(0.0), f(s), f(2 * s), s)
    assert abs(d - 0.43 * s) < 1e-14


@pytest.mark.parametrize("v", [[-0.3, 0.2, -0.05, 0.05, 0.15, 0.25],   # v1*v2 < 0: a kink behind the crossing
                               [-0.3, -0.2, -0.05, 0.05, -0.15, 0.25]])  # v3*v4 < 0: a kink beyond it
def test_kink_falls_back_to_quadratic_eno(v):
    s = 0.1
    v = np.array(v)
    assert po.six_distance(*v, s) == po.eno_quadratic_distance(v[1], v[2], v[3], v[4], s)


# --------------------------------------------------------------------------
# reinitialization (re_step.cu, WENO5RK3Reinitialization.m)
# --------------------------------------------------------------------------
def _plane(g, n, c, scale=1.0):
    n = np.asarray(n, dtype=float) / np.linalg.norm(n)
    return scale * (n[0] * g.X + n[1] * g.Y + n[2] * g.Z - c)


@pytest.mark.parametrize("scale", [1.0, 2.0])
def test_re_step_on_planes(scale):
    g = cases.Grid(np.linspace(-1, 1, 20), np.linspace(-0.9, 0.9, 18), np.linspace(-1.1, 1.1, 16))
    F = _plane(g, (0.3, -0.5, 0.8), 0.07, scale)
    d6, md = po.set_distance(F, *g.spacing)
    dt = 0.3 * md
    step = po.re_step(F, F < 0, dt, d6, *g.spacing)
    # away from the box faces and the interface the WENO derivatives of a
    # linear function are exact: step = sign(F) (|grad F| - 1) dt (the update
    # F - step moves F toward the distance function from both sides)
    far = (np.abs(F) > 4 * g.ds * scale)
    inner = np.zeros_like(far)
    inner[3:-3, 3:-3, 3:-3] = True
    sel = far & inner
    want = np.sign(F) * (scale - 1.0) * dt
    assert sel.sum() > 100
    assert np.abs(step[sel] - want[sel]).max() < 1e-13


def test_weno5_reinit_converges_to_sphere_distance():
    errs = []
    for n in (24, 32):
        g = cases.linspace_grid(n)
        r = np.sqrt(g.X ** 2 + g.Y ** 2 + g.Z ** 2)
        F = r ** 2 - 0.36
        d6, md = po.set_distance(F, *g.spacing)
        Fr = po.weno5_rk3_reinit(F, 150, d6, md, *g.spacing)
        band = np.abs(Fr) < 1.5 * g.ds
        errs.append(np.abs(Fr - (r - 0.6))[band].max())
        assert np.array_equal(np.sign(Fr), np.sign(F))
    assert errs[0] < 2e-5 and errs[1] < errs[0] / 2.5, errs


@pytest.mark.parametrize("rk2,cfl", [(False, 0.5), (True, 0.3)])
def test_eno2_reinit_reduces_gradient_error(rk2, cfl):
    g = cases.linspace_grid(32)
    F = cases.star_phi0(g)
    band = np.abs(F) < 3 * g.dx

    def grad_err(f):
        gy, gx, gz = np.gradient(f, g.dy, g.dx, g.dz, axis=(2, 1, 0))
        return np.abs(np.sqrt(gx ** 2 + gy ** 2 + gz ** 2) - 1)[band].mean()

    Fr = po.eno2_reinit(F, 100, *g.spacing, rk2=rk2, cfl=cfl)
    assert grad_err(Fr) < 0.1 * grad_err(F)
    # the zero level set moves by a fraction of a cell at most
    d0 = po.prefilled_distances(F.shape, *g.spacing)
    po.boundary_location(d0, F, *g.spacing, order=2)
    d1 = po.prefilled_distances(F.shape, *g.spacing)
    po.boundary_location(d1, Fr, *g.spacing, order=2)
    same = (d0[0] < g.dx) & (d1[0] < g.dx)
    assert same.sum() > 50 and np.abs(d0[0] - d1[0])[same].max() < 0.2 * g.dx


# --------------------------------------------------------------------------
# extension and surface redistance (WENO5RK3Extend.m, ...SurfaceRedistance.m)
# --------------------------------------------------------------------------
def _sphere(n, r0):
    g = cases.linspace_grid(n)
    r = np.sqrt(g.X ** 2 + g.Y ** 2 + g.Z ** 2)
    return g, r, np.ascontiguousarray(r - r0)


def test_extension_of_z_off_a_sphere():
    # src/7_0_paper/1_0_WENO/p08_test_extend.m:29-57: z extended off the sphere
    # r0 = 0.5 equals r0*sin(elevation) = r0*z/|x|
    errs = []
    for n in (24, 32):
        g, r, F = _sphere(n, 0.5)
        d6, md = po.set_distance(F, *g.spacing)
        G = po.weno5_rk3_extend(g.Z.copy(), 100, F, d6, md, *g.spacing)
        band = np.abs(F) < 1.5 * g.ds
        errs.append(np.abs(G - 0.5 * g.Z / r)[band].max())
    assert errs[0] < 5e-3 and errs[1] < errs[0], errs


def test_extension_keeps_constants():
    g, r, F = _sphere(20, 0.55)
    d6, md = po.set_distance(F, *g.spacing)
    C = np.full_like(F, 3.25)
    # exact up to the rounding of the quintic interface interpolant, whose
    # monomial form (shared_utilities.hpp:57-103) reproduces a constant only
    # to ~1e-13 relative
    assert np.abs(po.weno5_rk3_extend(C, 10, F, d6, md, *g.spacing) - 3.25).max() < 1e-11


def test_surface_redistance_gives_geodesic_distance():
    # src/7_0_paper/1_0_WENO/p11_test_surface_redistance_table.m:23-82:
    # A = exp(z) - 1 on the sphere r0 = 0.6 becomes r0 * elevation
    g, r, F = _sphere(28, 0.6)
    A = np.exp(g.Z) - 1
    An = po.surface_redistance(A, F, 20, 10, *g.spacing)
    elev = np.arcsin(np.clip(g.Z / np.maximum(r, 1e-12), -1, 1))
    band = (np.abs(F) < 1.5 * g.ds) & (np.abs(elev) < 1.2)
    err = np.abs(An - 0.6 * elev)[band]
    err0 = np.abs(A - 0.6 * elev)[band]
    assert err.max() < 0.15 * err0.max() and err.mean() < 5e-3, (err.max(), err.mean(), err0.max())


@pytest.mark.parametrize("profile", ["linear", "exp"])
def test_surface_redistance_keeps_the_zero_set(profile):
    # A = z - z0 (or exp(z) - exp(z0)) with z0 between nodes: on the sphere
    # r0 = 0.6 the geodesic distance to the circle A = 0 is r0 (elev - elev0).
    # Next to that circle the result is accurate to ~1e-3 only when A's
    # reinitialization uses A's own subcell distances and mask
    # (WENO5RK3ClosetPointSurfaceRedistance.m:30-52).
    g, r, F = _sphere(28, 0.6)
    z0 = g.zv[14] + 0.37 * g.dz
    A = g.Z - z0 if profile == "linear" else np.exp(g.Z) - np.exp(z0)
    An = po.surface_redistance(np.ascontiguousarray(A), F, 20, 10, *g.spacing)
    elev = np.arcsin(np.clip(g.Z / np.maximum(r, 1e-12), -1, 1))
    want = 0.6 * (elev - np.arcsin(z0 / 0.6))
    near = (np.abs(F) < 1.5 * g.ds) & (np.abs(want) < 2 * g.dz)
    assert near.sum() > 100
    err = np.abs(An - want)[near]
    assert err.max() < 3e-3 and err.mean() < 1e-3, (err.max(), err.mean())


# --------------------------------------------------------------------------
# advection (advection_step.cu + the Shu-Osher RK3 of the wrappers)
# --------------------------------------------------------------------------
def test_advection_translates_a_sphere():
    g, r, F = _sphere(32, 0.45)
    u = [np.full_like(F, 0.3), np.full_like(F, -0.2), np.full_like(F, 0.1)]
    T, steps = 0.5, 40
    Fa = po.advect_rk3(F, steps, T / steps, u, *g.spacing)
    ex = np.sqrt((g.X - 0.15) ** 2 + (g.Y + 0.1) ** 2 + (g.Z - 0.05) ** 2) - 0.45
    band = np.abs(ex) < 2 * g.ds
    assert np.abs(Fa - ex)[band].max() < 2e-3


def test_rk3_is_third_order_in_time():
    # same grid, dt halved twice: successive differences shrink ~8x (measured
    # away from the clamped box faces, whose kinks are not smooth in time)
    g = cases.linspace_grid(24)
    F = np.ascontiguousarray(np.sin(2 * g.X) * np.cos(g.Y) + 0.5 * g.Z)
    u = [np.full(g.shape, 0.7), np.full(g.shape, -0.4), np.full(g.shape, 0.3)]
    T = 0.2
    sols = [po.advect_rk3(F, k, T / k, u, *g.spacing) for k in (4, 8, 16)]
    inner = (slice(6, -6),) * 3
    d1 = np.abs(sols[0] - sols[1])[inner].max()
    d2 = np.abs(sols[1] - sols[2])[inner].max()
    assert 6.5 < d1 / d2 < 9.5, d1 / d2


# --------------------------------------------------------------------------
# area and volume
# --------------------------------------------------------------------------
def test_towers_area_volume_converge():
    # R/4_0_Dirac_Delta_and_Heavside_function/1_0_Towers_primal_function_method/test_error.m
    errs = []
    for n in (32, 48):
        g, r, F = _sphere(n, 0.6)
        a, v = po.area_volume_towers(F, *g.spacing)
        errs.append((abs(a / (4 * math.pi * 0.36) - 1), abs(v / (4 * math.pi * 0.216 / 3) - 1)))
    # second order in area; the volume error is already at the 1e-4 level and
    # oscillates with the grid alignment
    assert errs[1][0] < 1e-3 and errs[1][0] < errs[0][0] / 1.8, errs
    assert errs[0][1] < 2e-4 and errs[1][1] < 2e-4, errs


def test_towers_integrals_of_a_scaled_distance():
    # the Towers delta of F integrates to the area weighted by 1/|grad F|
    # (int delta(F) dV = int_Gamma dS/|grad F|): half the area for F = 2 d, with
    # contributions from points with d in [Ds/2, Ds) that only the |F| < 2Ds
    # mask of SDF3.m:75 keeps; H and the volume do not depend on the scale
    g, r, _ = _sphere(48, 0.6)
    F = np.ascontiguousarray(2.0 * (r - 0.6))
    a, v = po.area_volume_towers(F, *g.spacing)
    assert abs(a / (0.5 * 4 * math.pi * 0.36) - 1) < 1e-3
    assert abs(v / (4 * math.pi * 0.216 / 3) - 1) < 2e-4


def test_cosine_area_volume_match_closed_forms():
    # continuum integrals of the cosine-smoothed delta/H of half-width eps for
    # the exact distance of a sphere of radius R (derived, not in the reference):
    #   area = 4 pi (R^2 + eps^2 (1/3 - 2/pi^2)),  vol = 4 pi R^3 / 3 + 8 pi R eps^2 (1/6 - 1/pi^2)
    R = 0.6
    rel = []
    for n in (40, 64):
        g, r, F = _sphere(n, R)
        eps = 1.5 * g.dx
        a, v = po.area_volume_cosine(F, *g.spacing, eps)
        A = 4 * math.pi * (R ** 2 + eps ** 2 * (1 / 3 - 2 / math.pi ** 2))
        V = 4 * math.pi * R ** 3 / 3 + 8 * math.pi * R * eps ** 2 * (1 / 6 - 1 / math.pi ** 2)
        rel.append((abs(a / A - 1), abs(v / V - 1)))
    assert rel[1][0] < 2e-3 and rel[1][1] < 1e-4, rel


def test_compensated_sums_against_fsum():
    g = cases.Grid(np.linspace(-1, 1, 30), np.linspace(-1, 1, 26), np.linspace(-1, 1, 22))
    F = RNG.normal(size=g.shape) * 0.01  # every point inside the band but three
    # two huge values (outside the band) give the points around them deltas
    # some 1e15 times the others: summed in plain order the small terms that
    # follow lose their low bits (about 2e-13 of the total)
    F[5, 7, 8], F[5, 7, 10] = 1e14, -1e14
    F = np.ascontiguousarray(F)
    eps = 0.05
    a, v = po.area_volume_cosine(F, *g.spacing, eps)
    # recompute the summands with numpy (periodic central differences) and sum exactly
    fx = (np.roll(F, -1, axis=1) - np.roll(F, 1, axis=1)) / (2 * g.dx)
    fy = (np.roll(F, -1, axis=2) - np.roll(F, 1, axis=2)) / (2 * g.dy)
    fz = (np.roll(F, -1, axis=0) - np.roll(F, 1, axis=0)) / (2 * g.dz)
    gm = np.sqrt(fx ** 2 + fy ** 2 + fz ** 2)
    delta = gm * (1 + np.cos(math.pi * F / eps)) / eps / 2
    heav = (1 + F / eps + np.sin(math.pi * F / eps) / math.pi) / 2
    ev = g.dx * g.dy * g.dz
    assert abs(a - math.fsum(delta.ravel()) * ev) <= 1e-14 * abs(a)
    assert abs(v - math.fsum((1 - heav).ravel()) * ev) <= 1e-14 * abs(v)


# --------------------------------------------------------------------------
# WENO-RK3 / ENO-RK2 variants (1_1_WENORK3_Reinitialization, 2_0_ENORK2_Extend,
# 2_1_WENORK3_Extend and their +SD/@SDF3 wrappers)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("trial", range(6))
def test_cubic_distance_finds_the_root_of_a_cubic(trial):
    # cubic through (0,v0),(s,v1),(2s,v2),(3s,v3) with its only root in
    # (s, 2s): Newton converges to it (weno_rk3_reinitialization.cu:74-117)
    s = RNG.uniform(0.02, 0.2)
    root = RNG.uniform(1.05, 1.95) * s
    a, k = RNG.uniform(0.2, 3.0), RNG.choice([-1.0, 1.0]) * RNG.uniform(0.5, 2.0)
    f = lambda t: k * (t - root) * ((t - root) ** 2 + a * s * s)  # noqa: E731  (monotone: one real root)
    xs = np.arange(4) * s
    v = f(xs)
    assert v[0] * v[1] >= 0 and v[2] * v[3] >= 0 and v[1] * v[2] < 0
    assert abs(po.cubic_distance(*v, s) - (root - s)) < 1e-12 * s


def test_cubic_distance_exact_for_linear_data():
    s = 0.05
    v = 0.12 - 1.9 * np.arange(4) * s  # root at 0.12/1.9 in (s, 2s)
    assert abs(po.cubic_distance(*v, s) - (0.12 / 1.9 - s)) < 1e-15


@pytest.mark.parametrize("v", [[-0.2, 0.05, -0.05, -0.1],  # v0*v1 < 0: a kink behind the crossing
                               [0.2, 0.05, -0.05, 0.1]])   # v2*v3 < 0: a kink beyond it
def test_cubic_distance_kink_falls_back_to_quadratic_eno(v):
    s = 0.1
    v = np.array(v)
    assert po.cubic_distance(*v, s) == po.eno_quadratic_distance(*v, s)


def _nonuniform_stencil(h, r1, l1):
    """x3m..x3 of select_stencil (weno_rk3_reinitialization.cu:149-184)"""
    right = [r1, h, 2 * h] if r1 < h else [h, 2 * h, 3 * h]
    left = [-l1, -h, -2 * h] if l1 < h else [-h, -2 * h, -3 * h]
    return np.array(left[::-1] + [0.0] + right)


@pytest.mark.parametrize("trial", range(6))
def test_eno_cubic_derivative_exact_for_cubics(trial):
    # ENO_cubic_derivative (weno_rk3_reinitialization.cu:190-222) is the
    # derivative of a Newton cubic: exact for cubic data when the second
    # divided differences keep one sign (min_mod keeps the quadratic term)
    b, d = RNG.normal(size=2)
    c = 4.0 + abs(RNG.normal())  # p'' = 2c + 6dx > 0 on the stencil
    f = lambda t: 0.3 + b * t + c * t ** 2 + d * t ** 3  # noqa: E731
    h = 0.1
    r1, l1 = RNG.uniform(0.05, 0.95, size=2) * h
    for rr, ll in ((r1, h), (h, l1), (r1, l1), (h, h)):
        x = _nonuniform_stencil(h, rr, ll)
        df, db = po.eno_cubic_pair(x, f(x))
        assert abs(df - b) < 1e-12 and abs(db - b) < 1e-12


def test_wenork3_derivative_boundary_branches():
    h = 0.1
    x = (np.arange(7) - 3) * h
    p = 0.4 - 1.3 * x + 0.2 * x ** 2
    # no sign change: the uniform WENO5 pair of re_step.cu:11-37
    assert po.wenork3_derivative_boundary(p, h, h, h, rule=0) == po.weno_pair(p, h)
    # a sign change beside the node: cubic ENO with value 0 at the crossing
    q = x - 0.03
    r1 = 0.03
    df, db = po.wenork3_derivative_boundary(q, h, h, h, rule=0)
    xs = _nonuniform_stencil(h, h, h)
    assert (df, db) == po.eno_cubic_pair(xs, q)
    df, db = po.wenork3_derivative_boundary(q, r1, h, h, rule=1, v_fore=0.0)
    xs = _nonuniform_stencil(h, r1, h)
    hs = np.array([q[0], q[1], q[2], q[3], 0.0, q[4], q[5]])
    assert (df, db) == po.eno_cubic_pair(xs, hs)
    assert abs(df - 1.0) < 1e-13 and abs(db - 1.0) < 1e-13


def _axis_grid():
    return cases.Grid(np.linspace(-1, 1, 18), np.linspace(-0.9, 0.9, 16), np.linspace(-1.1, 1.1, 14))


@pytest.mark.parametrize("scheme,order,deg", [("enork2", 2, 2), ("wenork3", 3, 3)])
def test_variant_interface_values_exact_for_polynomials(scheme, order, deg):
    # boundary_interpolate: ENO2 quadratic (enork2_extend.cu:144-170) and
    # cubic (wenork3_extend.cu:310-329) interpolants reproduce polynomials of
    # their degree at the crossing
    g = _axis_grid()
    F = g.X - 0.013  # crossings between columns along x only
    d6 = po.prefilled_distances(F.shape, *g.spacing)
    po.boundary_location(d6, F, *g.spacing, order=order)
    C = 0.2 + 0.7 * g.X - 0.4 * g.X ** 2 + (0.3 * g.X ** 3 if deg == 3 else 0.0) + g.Y
    cp = po.variant_boundary_interpolate(d6, C, *g.spacing, scheme)
    x_at = lambda xs: 0.2 + 0.7 * xs - 0.4 * xs ** 2 + (0.3 * xs ** 3 if deg == 3 else 0.0)  # noqa: E731
    sel = d6[0] < g.dx
    assert sel.sum() > 50
    want = x_at(g.X[sel] + d6[0][sel]) + g.Y[sel]
    assert np.abs(cp[0][sel] - want).max() < 1e-12
    sel = d6[1] < g.dx
    want = x_at(g.X[sel] - d6[1][sel]) + g.Y[sel]
    assert np.abs(cp[1][sel] - want).max() < 1e-12
    assert np.all(cp[2] == 0) and np.all(cp[4] == 0)


@pytest.mark.parametrize("scheme", ["enork2", "wenork3"])
def test_variant_extension_keeps_fields_constant_along_normals(scheme):
    # F = x - x0: v = sign(F) e_x; a C that does not depend on x has zero
    # x-derivatives, so every extension step is an exact zero.  Heun's
    # (C + C - 0)/2 returns C exactly; the RK3 combinations 3/4 C + 1/4 C and
    # 1/3 C + 2/3 C round to within an ulp or two of C.
    g = _axis_grid()
    F = g.X - 0.013
    C = 0.3 + np.sin(2.0 * g.Y) * np.cos(g.Z)
    out = po.variant_extend(C, 8, F, *g.spacing, scheme)
    if scheme == "enork2":
        assert np.array_equal(out, C)
    else:
        assert np.abs(out - C).max() <= 4 * np.spacing(np.abs(C).max())


@pytest.mark.parametrize("scheme", ["enork2", "wenork3"])
def test_variant_extension_of_z_off_a_sphere(scheme):
    # as test_extension_of_z_off_a_sphere, with the ENO2 / WENO-RK3 schemes
    errs = []
    for n in (24, 32):
        g, r, F = _sphere(n, 0.5)
        G = po.variant_extend(g.Z.copy(), 100, F, *g.spacing, scheme)
        band = np.abs(F) < 1.5 * g.ds
        errs.append(np.abs(G - 0.5 * g.Z / r)[band].max())
    assert errs[0] < 2e-2 and errs[1] < errs[0], errs


def test_wenork3_reinit_keeps_planes_and_converges_on_spheres():
    g = _axis_grid()
    F = g.X - 0.013
    out = po.wenork3_reinit(F, 5, *g.spacing)
    assert np.abs(out - F).max() < 1e-13
    errs = []
    for n in (24, 32):
        gs = cases.linspace_grid(n)
        r = np.sqrt(gs.X ** 2 + gs.Y ** 2 + gs.Z ** 2)
        Fs = po.wenork3_reinit(r ** 2 - 0.36, 100, *gs.spacing)
        band = np.abs(r - 0.6) < 2 * gs.ds
        errs.append(np.abs(Fs - (r - 0.6))[band].max())
    assert errs[0] < 5e-3 and errs[1] < errs[0], errs


# --------------------------------------------------------------------------
# auxiliary-field geometry (4_0_Calculus_ToolBox/auxi_set_calculus_toolbox.cu,
# +SD/@SDF3/GPUAsetCalculusToolBox.m)
# --------------------------------------------------------------------------
_CF = np.array([0.07, -0.04, 0.11])
_HA = np.array([[0.6, -0.25, 0.0], [-0.25, 0.4, 0.15], [0.0, 0.15, 0.2]])
_BA = np.array([0.1, -0.2, 1.0])


def _quadratic_pair(g):
    """F = |x-c|^2/2 - 0.2 and A = x.HA.x/2 + b.x: central differences of
    quadratics are exact (equal spacings make the cross terms exact too)"""
    P = np.stack([g.X, g.Y, g.Z], axis=-1)
    F = 0.5 * ((P - _CF) ** 2).sum(-1) - 0.2
    A = 0.5 * np.einsum("...i,ij,...j->...", P, _HA, P) + P @ _BA
    return P, np.ascontiguousarray(F), np.ascontiguousarray(A)


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _frame(P):
    gF = P - _CF
    gA = P @ _HA + _BA
    N = _unit(gF)
    t = _unit(np.cross(gF, gA))
    return N, t, np.cross(t, N)


def test_aux_toolbox_frame_and_curvatures_match_the_analytic_curve():
    # t = grad F x grad A/|.|, n = t x N; geodesic curvature n.dt/ds, normal
    # curvature -t.dN/ds, geodesic torsion -t.D_n N and -n.D_n N, all by
    # central differences of the analytic frame along the curve (eps = 1e-5)
    g = cases.linspace_grid(20)
    P, F, A = _quadratic_pair(g)
    ft = po.set_calculus_toolbox(F, np.maximum(F, 0), *g.spacing, g.ds)
    fin = [ft[k] for k in po.TOOLBOX_FIELDS[:13]]
    out = po.auxi_set_calculus_toolbox(A, np.maximum(A, 0), fin, *g.spacing, g.ds)
    inner = (slice(1, -1),) * 3
    gA = P @ _HA + _BA
    C = np.cross(P - _CF, gA)
    sel = np.zeros(F.shape, bool)
    sel[inner] = True
    sel &= np.linalg.norm(C, axis=-1) > 0.05
    assert sel.sum() > 2000
    for k, name in enumerate(("Ax", "Ay", "Az")):
        assert np.abs(out[name] - gA[..., k])[sel].max() < 1e-12
    for name, v in zip(("Axx", "Ayy", "Azz", "Axy", "Ayz", "Azx"),
                       (_HA[0, 0], _HA[1, 1], _HA[2, 2], _HA[0, 1], _HA[1, 2], _HA[2, 0])):
        assert np.abs(out[name] - v)[sel].max() < 1e-10, name
    assert np.abs(out["ALaplacian"] - np.trace(_HA))[sel].max() < 1e-10
    for k, name in enumerate(("Cx", "Cy", "Cz")):
        assert np.abs(out[name] - C[..., k])[sel].max() < 1e-12
    assert np.abs(out["NormCrossAF"] - np.linalg.norm(C, axis=-1))[sel].max() < 1e-12
    N, t, n = _frame(P)
    for k in range(3):
        assert np.abs(out["t" + "xyz"[k]] - t[..., k])[sel].max() < 1e-12
        assert np.abs(out["n" + "xyz"[k]] - n[..., k])[sel].max() < 1e-12
    eps = 1e-5

    def along(fn, d):
        return (fn(P + eps * d) - fn(P - eps * d)) / (2 * eps)

    dt = along(lambda Q: _frame(Q)[1], t)
    dNt = along(lambda Q: _frame(Q)[0], t)
    dNn = along(lambda Q: _frame(Q)[0], n)
    checks = {
        "GeodesicCurvature": (n * dt).sum(-1),
        "NormalCurvature": -(t * dNt).sum(-1),
        "GeodesicTorsion": -(t * dNn).sum(-1),
        "BPerpendicular": -(n * dNn).sum(-1),
    }
    for name, want in checks.items():
        err = np.abs(out[name] - want)[sel].max()
        assert err < 1e-7 * max(1.0, np.abs(want[sel]).max()), (name, err)
    # on the sphere-like F: normal curvature -1/|x-c| for every tangent direction
    assert np.abs(out["NormalCurvature"] + 1 / np.linalg.norm(P - _CF, axis=-1))[sel].max() < 1e-9


def test_aux_heaviside_and_delta_away_from_the_aux_surface():
    # Towers H/delta of A from max(A,0): 1/0 where the 7-point stencil is
    # on the positive side, exactly 0/0 on the negative side
    g = cases.linspace_grid(16)
    A = np.ascontiguousarray(0.8 * g.X - 0.3 * g.Y + 0.5 * g.Z + 0.05)
    Fq = np.ascontiguousarray(g.X ** 2 + g.Y ** 2 + g.Z ** 2 - 0.5)
    ft = po.set_calculus_toolbox(Fq, np.maximum(Fq, 0), *g.spacing, g.ds)
    out = po.auxi_set_calculus_toolbox(A, np.maximum(A, 0), [ft[k] for k in po.TOOLBOX_FIELDS[:13]], *g.spacing,
                                       g.ds)
    h = 1.01 * max(0.8 * g.dx, 0.3 * g.dy, 0.5 * g.dz) * 3
    inner = np.zeros(A.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    pos, neg = inner & (A > h), inner & (A < -h)
    assert pos.sum() > 300 and neg.sum() > 300
    assert np.abs(out["AHeaviside"][pos] - 1).max() < 1e-13
    assert np.abs(out["ADiracDelta"][pos]).max() < 1e-12
    assert np.all(out["AHeaviside"][neg] == 0) and np.all(out["ADiracDelta"][neg] == 0)


def _roll(f, shift, axis):
    # circshift on the MATLAB dims (rows=y, cols=x, pges=z) of a (pges, cols, rows) array
    return np.roll(f, shift, axis={"x": 1, "y": 2, "z": 0}[axis])


def test_aux_toolbox_tail_is_gd3_periodic_differencing():
    # AFDiracDelta / ADiracDeltaDn of GPUAsetCalculusToolBox.m:57-63 written
    # with np.roll as @GD3's circshift maps (GD3.m:112-117): Field(oXo) is the
    # right neighbour, wrapping at the box face
    g = cases.small_grids()["star"][0]
    F = cases.star_phi0(g)
    A = np.ascontiguousarray(0.7 * g.Z + 0.2 * g.X * g.Y - 0.1)
    ft = po.set_calculus_toolbox(F, np.maximum(F, 0), *g.spacing, g.ds)
    out = po.aux_calculus_toolbox(A, ft, *g.spacing, g.ds)
    assert np.array_equal(out["AHPrimal"], np.maximum(A, 0))
    dx, dy, dz = g.spacing

    def grad(f):
        return [(_roll(f, -1, "x") - _roll(f, 1, "x")) / (2 * dx), (_roll(f, -1, "y") - _roll(f, 1, "y")) / (2 * dy),
                (_roll(f, -1, "z") - _roll(f, 1, "z")) / (2 * dz)]

    def lap(f):
        return ((_roll(f, -1, "x") - 2 * f + _roll(f, 1, "x")) / (dx * dx)
                + (_roll(f, -1, "y") - 2 * f + _roll(f, 1, "y")) / (dy * dy)
                + (_roll(f, -1, "z") - 2 * f + _roll(f, 1, "z")) / (dz * dz))

    h1 = grad(out["AHeaviside"])
    h2 = grad(ft["Heaviside"])
    ch = [h2[1] * h1[2] - h2[2] * h1[1], h2[2] * h1[0] - h2[0] * h1[2], h2[0] * h1[1] - h2[1] * h1[0]]
    afdd = (out["Cx"] * ch[0] + out["Cy"] * ch[1] + out["Cz"] * ch[2]) / out["NormCrossAF"] ** 2
    addn = (lap(out["AHeaviside"]) - out["ADiracDelta"] * lap(A)) / out["AGradMag"]
    assert np.array_equal(out["AFDiracDelta"], afdd)
    assert np.array_equal(out["ADiracDeltaDn"], addn)
    # the wrap is what makes the faces differ from a clamped stencil
    assert np.abs(lap(A) - np.pad(lap(A)[1:-1, 1:-1, 1:-1], 1, mode="edge")).max() > 1.0


def test_aux_dirac_delta_integrates_to_the_curve_length():
    # AFDiracDelta ~ delta(F) delta(A): with F = |x| - R and A = z the
    # integral is the length 2 pi R of the equator (|grad F x grad A| = 1 there)
    R = 0.55
    lens = []
    for n in (40, 56):
        g = cases.linspace_grid(n)
        F = np.ascontiguousarray(np.sqrt(g.X ** 2 + g.Y ** 2 + g.Z ** 2) - R)
        A = np.ascontiguousarray(g.Z + 0.013)
        ft = po.set_calculus_toolbox(F, np.maximum(F, 0), *g.spacing, g.ds)
        out = po.aux_calculus_toolbox(A, ft, *g.spacing, g.ds)
        lens.append(out["AFDiracDelta"].sum() * g.dx * g.dy * g.dz)
    want = 2 * np.pi * np.sqrt(R ** 2 - 0.013 ** 2)
    errs = [abs(v - want) / want for v in lens]
    assert errs[0] < 3e-3 and errs[1] < errs[0], (lens, want)
