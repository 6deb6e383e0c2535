"""Known answers for the oracle's kernel-level entry points, on fields whose
results the mathematics fixes: lines with a kink at a node (each one-sided
WENO5 / ENO2 derivative then sees only its own side's slope), linear and
quadratic fields (central differences, the quintic interpolant and WENO5 are
exact on them), and planes whose crossings are known.

They pin what polynomial exactness alone leaves open: which one-sided
derivative a kernel picks (upwind_normal.cu:10, enork2_extend.cu:98,
wenork3_extend.cu:198; extend_step.cu:121-123; advection_step.cu:51-53),
which output slot and axis a value goes to, the branch between the uniform
and the nonuniform stencil, frozen points, and the toolbox's 4*ds^2
cross-derivative denominator (set_calculus_toolbox.cu:64-66).
"""
import numpy as np
import pytest

import cases
from oracle import pyoracle as po

SP = (0.1, 0.13, 0.07)  # dx, dy, dz: distinct, so a swapped spacing shows
N = 13
K = 6  # the kink node on the varied axis


def _grid():
    dx, dy, dz = SP
    t = np.arange(N) - K
    return cases.Grid(t * dx, t * dy, t * dz)


def _coord(g, axis):
    """coordinates along axis 0 = x (cols), 1 = y (rows), 2 = z (pages)"""
    return (g.X, g.Y, g.Z)[axis]


def _kinked(g, axis, right, left, c=5.0):
    """c + right*t for t >= 0, c + left*t for t < 0 along `axis` (kink at
    node K), constant along the other axes; positive everywhere"""
    t = _coord(g, axis)
    return np.ascontiguousarray(c + np.where(t >= 0, right * t, left * t))


def _at_kink(a, axis):
    """the values of a 3-D field on the plane through the kink node"""
    sl = [slice(None)] * 3
    sl[{0: 1, 1: 2, 2: 0}[axis]] = K
    return a[tuple(sl)]


def _spacing_distances(F):
    return po.prefilled_distances(F.shape, *SP)


# (right, left) slopes: p5 = c + right*h, p3 = c - left*h, so |p5| < |p3| iff
# right < -left; the expected normal is the right slope then, else the left
SIDES = [(-2.0, 1.0), (2.0, -1.0), (1.0, -3.0), (-1.0, 3.0)]


def _expected_side(right, left):
    return right if right < -left else left


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("right,left", SIDES)
def test_upwind_normal_takes_the_side_with_the_smaller_neighbour(axis, right, left):
    # upwind_normal.cu:10: fore derivative iff |p5| < |p3|.  At a kink node the
    # fore WENO5 derivative has one smooth candidate stencil (slope `right`)
    # and the back one likewise (slope `left`); their weights are 1 up to
    # (1e-6 / S)^2, so each returns its side's slope to ~1e-11.
    g = _grid()
    F = _kinked(g, axis, right, left)
    n = po.upwind_normal(F, _spacing_distances(F), *SP)
    want = _expected_side(right, left)
    assert np.abs(_at_kink(n[axis], axis) - want).max() < 1e-9
    for other in {0, 1, 2} - {axis}:
        assert np.abs(n[other]).max() < 1e-12


@pytest.mark.parametrize("scheme", ["enork2", "wenork3"])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("right,left", SIDES)
def test_variant_upwind_normals_take_the_same_side(scheme, axis, right, left):
    # enork2_extend.cu:98 (|vr| < |vl|, vr = p5 without a crossing) and
    # wenork3_extend.cu:198 (|p5| < |p3|).  ENO2 with min_mod is exact on
    # each side of the kink.
    g = _grid()
    F = _kinked(g, axis, right, left)
    n = po.variant_upwind_normal(F, _spacing_distances(F), *SP, scheme)
    want = _expected_side(right, left)
    assert np.abs(_at_kink(n[axis], axis) - want).max() < 1e-9
    for other in {0, 1, 2} - {axis}:
        assert np.abs(n[other]).max() < 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_weno_derivative_slots(axis):
    # nonoscillating_interpolant.cu:59-122 writes back_x, fore_x, back_y, ...
    g = _grid()
    F = _kinked(g, axis, 1.5, -0.5)
    out = po.weno_derivative(F, *SP)
    assert np.abs(_at_kink(out[2 * axis], axis) + 0.5).max() < 1e-9  # back: left slope
    assert np.abs(_at_kink(out[2 * axis + 1], axis) - 1.5).max() < 1e-9  # fore: right slope
    for k in range(6):
        if k // 2 != axis:
            assert np.abs(out[k]).max() < 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("vel", [0.7, -0.7])
def test_advection_and_extension_steps_are_upwind(axis, vel):
    # min(0,v) D+ + max(0,v) D- (advection_step.cu:51-53, extend_step.cu:121-123):
    # information comes from the side the velocity comes from
    g = _grid()
    C = _kinked(g, axis, 1.5, -0.5)
    v3 = [np.zeros_like(C) for _ in range(3)]
    v3[axis][:] = vel
    want = vel * (-0.5 if vel > 0 else 1.5)
    a = po.advection_step(v3, C, *SP)
    assert np.abs(_at_kink(a, axis) - want).max() < 1e-9
    dt = np.full_like(C, 0.25)
    d6 = _spacing_distances(C)
    cp6 = [np.zeros_like(C) for _ in range(6)]
    s = po.extend_step(dt, C, v3, d6, cp6, *SP)
    assert np.abs(_at_kink(s, axis) - 0.25 * want).max() < 1e-9
    for scheme in ("enork2", "wenork3"):
        s = po.variant_extend_step(dt, C, v3, d6, cp6, *SP, scheme)
        assert np.abs(_at_kink(s, axis) - 0.25 * want).max() < 1e-9, scheme


def test_linear_fields_advect_and_extend_exactly():
    g = _grid()
    grad = (0.3, -1.1, 0.45)
    C = np.ascontiguousarray(0.2 + grad[0] * g.X + grad[1] * g.Y + grad[2] * g.Z)
    rng = np.random.default_rng(7)
    v3 = [np.ascontiguousarray(rng.uniform(-1, 1, C.shape)) for _ in range(3)]
    want = grad[0] * v3[0] + grad[1] * v3[1] + grad[2] * v3[2]
    # interior nodes: the clamped stencils at the faces see a kink
    inner = (slice(3, -3),) * 3
    assert np.abs(po.advection_step(v3, C, *SP)[inner] - want[inner]).max() < 1e-12
    dt = np.ascontiguousarray(rng.uniform(0.1, 0.3, C.shape))
    d6 = _spacing_distances(C)
    cp6 = [np.zeros_like(C) for _ in range(6)]
    s = po.extend_step(dt, C, v3, d6, cp6, *SP)
    assert np.abs(s[inner] - (dt * want)[inner]).max() < 1e-12


def _plane_crossing(g, axis, t0):
    """F = t - t0 along `axis` and its subcell distances (linear data: the
    quintic root is exact)"""
    F = np.ascontiguousarray(_coord(g, axis) - t0)
    d6 = po.prefilled_distances(F.shape, *SP)
    po.boundary_location(d6, F, *SP)
    return F, d6


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_boundary_location_slots_and_roots(axis):
    # boundary_location.cu:84-132: the r-side distance at the node before the
    # crossing, the l-side distance at the node after it; the others keep the
    # pre-filled spacing
    g = _grid()
    h = SP[axis]
    t0 = 0.37 * h
    F, d6 = _plane_crossing(g, axis, t0)
    r, l = d6[2 * axis], d6[2 * axis + 1]
    t = _coord(g, axis)
    before = np.isclose(t, 0.0)
    after = np.isclose(t, h)
    assert np.abs(r[before] - t0).max() < 1e-15 and np.abs(l[after] - (h - t0)).max() < 1e-15
    assert np.all(r[~before] == h) and np.all(l[~after] == h)
    for other in {0, 1, 2} - {axis}:
        assert np.all(d6[2 * other] == SP[other]) and np.all(d6[2 * other + 1] == SP[other])


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_boundary_interpolant_is_exact_for_quintics(axis):
    # boundary_interpolant.cu:77-141: the quintic through six nodes, evaluated
    # at the crossing; fore value at the r-side node, back value at the l-side
    # node, zero elsewhere
    g = _grid()
    h = SP[axis]
    t0 = 0.37 * h
    F, d6 = _plane_crossing(g, axis, t0)
    t = _coord(g, axis)
    poly = lambda s: 0.3 - 0.8 * s + 1.1 * s ** 2 + 0.5 * s ** 3 - 0.9 * s ** 4 + 0.7 * s ** 5  # noqa: E731
    other = (g.X, g.Y, g.Z)[(axis + 1) % 3]
    C = np.ascontiguousarray(poly(t) + other)
    cp = po.boundary_interpolant(d6, C, *SP)
    want = poly(t0) + other
    before, after = np.isclose(t, 0.0), np.isclose(t, h)
    # the monomial form (shared_utilities.hpp:57-103) is ill-conditioned:
    # ~1e-11 relative
    assert np.abs(cp[2 * axis][before] - want[before]).max() < 1e-10
    assert np.abs(cp[2 * axis + 1][after] - want[after]).max() < 1e-10
    assert np.all(cp[2 * axis][~before] == 0) and np.all(cp[2 * axis + 1][~after] == 0)
    for k in range(6):
        if k // 2 != axis:
            assert np.all(cp[k] == 0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_extend_step_at_the_interface(axis):
    # extend_step.cu:10-125 with shared_utilities.cup:256-278: where a distance
    # is below the spacing the nonuniform stencil takes the interface value;
    # on a linear C with exact interface values the step stays exact.  A
    # point with a distance below 1e-6*dx is frozen (step 0).
    g = _grid()
    h = SP[axis]
    F, d6 = _plane_crossing(g, axis, 0.37 * h)
    grad = np.array([0.3, -1.1, 0.45])
    C = np.ascontiguousarray(0.2 + grad[0] * g.X + grad[1] * g.Y + grad[2] * g.Z)
    cp = po.boundary_interpolant(d6, C, *SP)
    v3 = [np.full_like(C, s) for s in (0.6, -0.4, 0.5)]
    dt = np.full_like(C, 0.2)
    s = po.extend_step(dt, C, v3, d6, cp, *SP)
    want = 0.2 * (grad[0] * 0.6 - grad[1] * 0.4 + grad[2] * 0.5)
    inner = (slice(3, -3),) * 3
    assert np.abs(s[inner] - want).max() < 1e-11
    # the interface branch: a planted distance of half a cell on the upwind
    # side (l for a positive velocity), with interface value 0 there, bends
    # that side's stencil, so the step differs from the uniform one
    side = 2 * axis + (1 if (0.6, -0.4, 0.5)[axis] > 0 else 0)
    node = [K, K, K]
    node[{0: 1, 1: 2, 2: 0}[axis]] = K + 2  # two nodes past the crossing: cp is 0 there
    node = tuple(node)
    d6b = [d.copy() for d in d6]
    d6b[side][node] = 0.5 * h
    s2 = po.extend_step(dt, C, v3, d6b, cp, *SP)
    assert abs(s2[node] - want) > 1e-3
    # frozen
    d6b[side][node] = 1e-7 * SP[0]
    s3 = po.extend_step(dt, C, v3, d6b, cp, *SP)
    assert s3[node] == 0.0 and s3[K, K, K] == s[K, K, K]


def test_set_calculus_toolbox_on_a_quadratic():
    # set_calculus_toolbox.cu:31-106 on F = x^T Q x / 2 + b.x + c: central
    # differences are exact on quadratics, except that the cross derivatives
    # are divided by 4 ds^2 where 4 dx dy would be exact (:64-66).  Mean and
    # Gaussian curvature of an implicit surface (R. Goldman, CAGD 22 (2005),
    # eq. 4.1 and 4.2, written with numpy's determinant and inverse):
    # -(|g|^2 tr H - g^T H g) / |g|^3 and g^T adj(H) g / |g|^4.
    g = _grid()
    dx, dy, dz = SP
    ds = (dx + dy + dz) / 3
    Q = np.array([[1.3, 0.4, -0.2], [0.4, -0.7, 0.3], [-0.2, 0.3, 0.9]])
    b = np.array([0.5, -0.2, 0.8])
    X = np.stack([g.X, g.Y, g.Z], axis=-1)
    F = np.ascontiguousarray(0.5 * np.einsum("...i,ij,...j->...", X, Q, X) + X @ b - 0.05)
    P = np.ascontiguousarray(np.maximum(F, 0.0))
    out = po.set_calculus_toolbox(F, P, dx, dy, dz, ds)
    inner = (slice(1, -1),) * 3
    grad = X @ Q + b
    scale = {(0, 1): dx * dy / ds ** 2, (1, 2): dy * dz / ds ** 2, (0, 2): dz * dx / ds ** 2}
    H = Q.copy()
    for (i, j), s in scale.items():
        H[i, j] = H[j, i] = Q[i, j] * s
    names = ["Fx", "Fy", "Fz"]
    for i in range(3):
        assert np.abs(out[names[i]][inner] - grad[..., i][inner]).max() < 1e-12
    gm = np.maximum(np.sqrt((grad ** 2).sum(-1)), 1e-14)
    assert np.abs(out["FGradMag"][inner] - gm[inner]).max() < 1e-12
    for i, nm in enumerate(["Nx", "Ny", "Nz"]):
        assert np.abs(out[nm][inner] - (grad[..., i] / gm)[inner]).max() < 1e-11
    for nm, (i, j) in zip(["Fxx", "Fyy", "Fzz", "Fxy", "Fyz", "Fzx"],
                          [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]):
        assert np.abs(out[nm][inner] - H[i, j]).max() < 1e-10, nm
    assert np.abs(out["FLaplacian"][inner] - np.trace(H)).max() < 1e-10
    gg = grad[inner]
    gn = np.sqrt((gg ** 2).sum(-1))
    ghg = np.einsum("...i,ij,...j->...", gg, H, gg)
    mean = -(gn ** 2 * np.trace(H) - ghg) / gn ** 3
    adj = np.linalg.det(H) * np.linalg.inv(H)
    gauss = np.einsum("...i,ij,...j->...", gg, adj, gg) / gn ** 4
    ok = gn > 0.2
    assert np.abs(out["MeanCurvature"][inner][ok] - mean[ok]).max() < 1e-9
    assert np.abs(out["GaussianCurvature"][inner][ok] - gauss[ok]).max() < 1e-9
    # Towers H and delta (:91-106): where max(F,0) = F on the whole stencil
    # H = 1 and delta = 0; where it is 0 there, both are 0
    pos = np.ones_like(F, dtype=bool)
    neg = np.ones_like(F, dtype=bool)
    for ax in range(3):
        for sh in (-1, 0, 1):
            Fs = np.roll(F, sh, axis=ax)
            pos &= Fs > 0
            neg &= Fs < 0
    pos[~np.pad(np.ones((N - 2,) * 3, bool), 1)] = False
    neg[~np.pad(np.ones((N - 2,) * 3, bool), 1)] = False
    assert pos.sum() > 100 and neg.sum() > 10
    assert np.abs(out["Heaviside"][pos] - 1).max() < 1e-12 and np.abs(out["DiracDelta"][pos]).max() < 1e-9
    assert np.all(out["Heaviside"][neg] == 0) and np.all(out["DiracDelta"][neg] == 0)


SCHEME_ORDER = {"weno5": 6, "eno2": 2, "wenork3": 3}


def _oblique_plane(g, scale, c=0.031, signs=(1, -1, 1)):
    n = np.array([0.48, 0.6, 0.64]) * np.asarray(signs, dtype=float)
    n = n / np.linalg.norm(n)
    return np.ascontiguousarray(scale * (n[0] * g.X + n[1] * g.Y + n[2] * g.Z - c))


@pytest.mark.parametrize("signs", [(1, -1, 1), (-1, 1, -1), (1, 1, 1), (-1, -1, -1)])
@pytest.mark.parametrize("scheme", ["weno5", "eno2", "wenork3"])
def test_re_step_exact_on_planes_including_interface_nodes(scheme, signs):
    # On linear data with exact crossing distances every one-sided derivative
    # is exact, also the subcell ones with the interface value 0 at the
    # crossing (re_step.cu:41-77, boundary_correction.cu:80-100,
    # weno_rk3_reinitialization.cu:149-222): step = sign(F)(|grad F| - 1) dt
    # at every node whose stencil stays inside the box.
    g = cases.Grid(np.linspace(-1, 1, 20), np.linspace(-0.9, 0.9, 18), np.linspace(-1.1, 1.1, 16))
    F = _oblique_plane(g, 2.0, signs=signs)
    d6, md = po.set_distance(F, *g.spacing, order=SCHEME_ORDER[scheme])
    assert md.min() > 1e-3 * g.dx  # no frozen node
    dt = 0.3 * md
    step = po.re_step(F, F < 0, dt, d6, *g.spacing, scheme=scheme)
    inner = np.zeros(F.shape, bool)
    inner[3:-3, 3:-3, 3:-3] = True
    sub = inner & (md < min(g.spacing))  # nodes next to the interface
    assert sub.sum() > 50
    want = np.sign(F) * (2.0 - 1.0) * dt
    assert np.abs(step - want)[inner].max() < 1e-12


@pytest.mark.parametrize("kind", ["weno5", "wenork3", "eno2", "extend", "extend_wenork3", "extend_enork2"])
def test_frozen_points(kind):
    # re_step.cu:290-300, weno_rk3_reinitialization.cu, extend_step.cu and
    # wenork3_extend.cu freeze a point when any of its six distances is below
    # 1e-6*dx (dx also for the y and z distances); the ENO2 re_step and the
    # ENO-RK2 extend_step have no such test.
    g = _grid()
    F = _oblique_plane(g, 1.0, c=5.0)  # no crossing in the box
    d6 = _spacing_distances(F)
    dt = np.full_like(F, 0.01)
    node = (K, K, K)
    zpu = 4  # the z distance: 1e-6*dz < planted < 1e-6*dx
    planted = 0.85e-6 * SP[0]
    assert 1e-6 * SP[2] < planted < 1e-6 * SP[0]

    def run(d):
        if kind in ("weno5", "wenork3", "eno2"):
            return po.re_step(F, F < 0, dt, d, *SP, scheme=kind)
        v3 = [np.full_like(F, s) for s in (0.6, -0.4, 0.5)]
        cp6 = [np.zeros_like(F) for _ in range(6)]
        if kind == "extend":
            return po.extend_step(dt, F, v3, d, cp6, *SP)
        return po.variant_extend_step(dt, F, v3, d, cp6, *SP, kind.split("_")[1])

    d = [a.copy() for a in d6]
    d[zpu][node] = planted
    s = run(d)
    if kind in ("eno2", "extend_enork2"):
        assert s[node] != 0.0
    else:
        assert s[node] == 0.0
        d[zpu][node] = 1.2e-6 * SP[0]
        assert run(d)[node] != 0.0


@pytest.mark.parametrize("order", [6, 2, 3])
def test_set_distance_min_dist(order):
    # setDistance.m:17-21: min_dist is the smallest of the six distances
    g = _grid()
    h = SP[2]  # the smallest spacing: both distances are below the others
    F, _ = _plane_crossing(g, 2, 0.37 * h)
    d6, md = po.set_distance(F, *SP, order=order)
    t = g.Z
    before, after = np.isclose(t, 0.0), np.isclose(t, h)
    assert np.abs(md[before] - 0.37 * h).max() < 1e-15
    assert np.abs(md[after] - 0.63 * h).max() < 1e-15
    assert np.all(md[~(before | after)] == min(SP))


def _far_field_grid():
    return cases.Grid(np.linspace(-1, 1, 40), np.linspace(-0.9, 0.9, 36), np.linspace(-1.1, 1.1, 32))


def _far(g, x_interface_index):
    """nodes at least 10 nodes from every box face and from the interface"""
    sel = np.zeros(g.shape, bool)
    sel[10:-10, 10:x_interface_index - 10, 10:-10] = True
    return sel


@pytest.mark.parametrize("scheme,cfl", [("weno5", 0.3), ("eno2_euler", 0.5), ("eno2_rk2", 0.3), ("wenork3", 0.3)])
def test_one_reinit_iteration_far_from_the_interface(scheme, cfl):
    # Far from the interface and the faces F = 2 (x - x0) stays linear through
    # every stage, so each stage's step is the constant -(2 - 1) dt inside
    # (mask F < 0) with dt = cfl * min(dx, dy, dz), and one RK3 / Heun / Euler
    # iteration moves F by exactly that (WENO5RK3Reinitialization.m:10-39,
    # ENORK2Reinitialization.m:22-40, WENORK3Reinitialization.m:27-57).
    g = _far_field_grid()
    x0 = 0.8
    F = np.ascontiguousarray(2.0 * (g.X - x0))
    ix = int(np.searchsorted(g.xv, x0))
    if scheme == "weno5":
        d6, md = po.set_distance(F, *g.spacing)
        out = po.weno5_rk3_reinit(F, 1, d6, md, *g.spacing)
    elif scheme == "wenork3":
        out = po.wenork3_reinit(F, 1, *g.spacing)
    else:
        out = po.eno2_reinit(F, 1, *g.spacing, rk2=scheme.endswith("rk2"), cfl=cfl)
    far = _far(g, ix)
    assert far.sum() > 500
    shift = cfl * min(g.spacing)
    assert np.abs((out - F)[far] - shift).max() < 1e-13


@pytest.mark.parametrize("scheme", ["weno5", "enork2", "wenork3"])
@pytest.mark.parametrize("scale", [1.0, 1e-9])
def test_one_extension_iteration_far_from_the_interface(scheme, scale):
    # F = scale (x - x0): the upwind normal is scale*e_x, the velocity
    # sign(F) n / max(|n|, 1e-6) (WENO5RK3Extend.m:23-31, ENORK2Extend.m,
    # WENORK3Extend.m) is -e_x on the far side for scale 1 and -1e-3 e_x for
    # scale 1e-9 (the floor).  On a linear C every stage's step is then the
    # constant dt v.grad C with dt = 0.3 min(dx, dy, dz), and one RK3 / Heun
    # iteration moves C by exactly -dt v.grad C.
    g = _far_field_grid()
    x0 = 0.8
    F = np.ascontiguousarray(scale * (g.X - x0))
    ix = int(np.searchsorted(g.xv, x0))
    C = np.ascontiguousarray(0.2 + 0.3 * g.X + 0.7 * g.Y - 0.4 * g.Z)
    if scheme == "weno5":
        d6, md = po.set_distance(F, *g.spacing)
        out = po.weno5_rk3_extend(C, 1, F, d6, md, *g.spacing)
    else:
        out = po.variant_extend(C, 1, F, *g.spacing, scheme)
    far = _far(g, ix)
    vx = -min(scale / 1e-6, 1.0)
    want = -0.3 * min(g.spacing) * vx * 0.3
    assert np.abs((out - C)[far] - want).max() < 1e-13 * max(1.0, abs(want) * 1e3)


def test_extension_velocity_is_zero_on_exact_zeros():
    # MATLAB sign(0) = 0 (WENO5RK3Extend.m:29-31): nodes where F is exactly 0
    # get no velocity; elsewhere sign(F) grad F/|grad F|
    g = _grid()
    F = np.ascontiguousarray(g.X - g.xv[K + 1])
    d6 = po.prefilled_distances(F.shape, *SP)
    po.boundary_location(d6, F, *SP)
    v = po.extension_velocity(F, d6, 1e-6, *SP)
    zero = F == 0
    assert zero.sum() == N * N
    for a in range(3):
        assert np.all(v[a][zero] == 0)
    inner = np.zeros(F.shape, bool)
    inner[3:-3, 3:-3, 3:-3] = True
    sel = inner & ~zero
    assert np.abs(v[0][sel] - np.sign(F[sel])).max() < 1e-12
    assert np.abs(v[1][sel]).max() < 1e-12 and np.abs(v[2][sel]).max() < 1e-12


def test_normal_velocity_of_a_linear_field():
    # u.N with N = grad F / |grad F| from central differences (exact on a
    # linear F)
    g = _grid()
    grad = np.array([0.3, -1.1, 0.45])
    F = np.ascontiguousarray(0.2 + grad[0] * g.X + grad[1] * g.Y + grad[2] * g.Z)
    rng = np.random.default_rng(3)
    v3 = [np.ascontiguousarray(rng.uniform(-1, 1, F.shape)) for _ in range(3)]
    n = grad / np.linalg.norm(grad)
    want = n[0] * v3[0] + n[1] * v3[1] + n[2] * v3[2]
    inner = (slice(1, -1),) * 3
    assert np.abs(po.normal_velocity(F, v3, *SP)[inner] - want[inner]).max() < 1e-13


@pytest.mark.parametrize("scheme", ["enork2", "wenork3"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_variant_extend_step_at_the_interface(scheme, axis):
    # enork2_extend.cu:56-78 / wenork3_extend.cu:167-190: next to a crossing
    # the one-sided derivatives take the interface value at the crossing;
    # with exact interface values of a linear C the step stays exact.
    g = _grid()
    h = SP[axis]
    F = np.ascontiguousarray(_coord(g, axis) - 0.37 * h)
    d6 = po.prefilled_distances(F.shape, *SP)
    po.boundary_location(d6, F, *SP, order=2 if scheme == "enork2" else 3)
    grad = np.array([0.3, -1.1, 0.45])
    C = np.ascontiguousarray(0.2 + grad[0] * g.X + grad[1] * g.Y + grad[2] * g.Z)
    cp = po.variant_boundary_interpolate(d6, C, *SP, scheme)
    v3 = [np.full_like(C, s) for s in (0.6, -0.4, 0.5)]
    dt = np.full_like(C, 0.2)
    s = po.variant_extend_step(dt, C, v3, d6, cp, *SP, scheme)
    want = 0.2 * (grad[0] * 0.6 - grad[1] * 0.4 + grad[2] * 0.5)
    inner = (slice(3, -3),) * 3
    assert np.abs(s[inner] - want).max() < 1e-11
    side = 2 * axis + (1 if (0.6, -0.4, 0.5)[axis] > 0 else 0)
    node = [K, K, K]
    node[{0: 1, 1: 2, 2: 0}[axis]] = K + 2
    node = tuple(node)
    d6b = [d.copy() for d in d6]
    d6b[side][node] = 0.5 * h
    assert abs(po.variant_extend_step(dt, C, v3, d6b, cp, *SP, scheme)[node] - want) > 1e-3


def _smooth(t):
    return np.sin(1.3 * t + 0.3) + 0.2 * t


def _smooth_slope(t):
    return 1.3 * np.cos(1.3 * t + 0.3) + 0.2


@pytest.mark.parametrize("nonuniform", [False, True])
def test_weno5_is_fifth_order_on_smooth_data(nonuniform):
    # Every candidate stencil is exact for cubics whatever the weights, so
    # polynomial exactness leaves the linear weights (0.1, 0.6, 0.3 of
    # re_step.cu:30-32; c0..c2 of the nonuniform form) open; fifth order on
    # smooth data needs them: halving h divides the error by ~32 (third order,
    # ~8, with other weights).
    errs = []
    for h in (0.04, 0.02, 0.01):
        x = (np.arange(7) - 3) * h
        if nonuniform:
            r1, l1 = 0.4 * h, 0.7 * h
            df, db = po.weno_derivative_boundary(_smooth(x), r1, l1, h, _smooth(r1), _smooth(-l1), rule=1)
        else:
            df, db = po.weno_pair(_smooth(x), h)
        errs.append(max(abs(df - _smooth_slope(0.0)), abs(db - _smooth_slope(0.0))))
    for a, b in zip(errs, errs[1:]):
        assert 24 < a / b < 40, errs


@pytest.mark.parametrize("order,degree", [(6, 5), (2, 2), (3, 3)])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_boundary_location_exact_for_polynomials(order, degree, axis):
    # the crossing distances of boundary_location (quintic + Brent),
    # boundary_correction (quadratic ENO) and the cubic boundary_correction
    # (Newton on the cubic) are exact for polynomials of their degree
    g = _grid()
    h = SP[axis]
    t = _coord(g, axis) / h  # node units
    root = 0.37
    extra = {5: (t + 4.1) * (t - 5.3) * (t + 6.2) * (t - 7.4), 2: (t + 3.3), 3: (t + 3.3) * (t - 4.6)}[degree]
    F = np.ascontiguousarray((t - root) * extra)
    d6 = po.prefilled_distances(F.shape, *SP)
    po.boundary_location(d6, F, *SP, order=order)
    before, after = np.isclose(t, 0.0), np.isclose(t, 1.0)
    assert np.abs(d6[2 * axis][before] - root * h).max() < 1e-13 * h
    assert np.abs(d6[2 * axis + 1][after] - (1 - root) * h).max() < 1e-13 * h


@pytest.mark.parametrize("scheme", ["weno5", "wenork3"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_re_step_interface_branch_follows_the_sign_test(scheme, axis):
    # re_step.cu:248-269 / weno_rk3_reinitialization.cu:228-247 take the
    # subcell stencil when a neighbour's value has the other sign (p3*p4 < 0 or
    # p4*p5 < 0), and then read the distances; without a sign change the
    # distances are not read.  A planted (wrong) l-distance therefore moves
    # the step of the node just past the crossing and leaves the node two past
    # it alone.
    g = _grid()
    h = SP[axis]
    F, d6 = _plane_crossing(g, axis, 0.37 * h)
    F = np.ascontiguousarray(F + 0.0)
    md = np.min(np.stack(d6), axis=0)
    dt = 0.3 * md
    base = po.re_step(F, F < 0, dt, d6, *SP, scheme=scheme)
    ax = {0: 1, 1: 2, 2: 0}[axis]
    for offset, moves in ((1, True), (2, False)):
        node = [K, K, K]
        node[ax] = K + offset
        node = tuple(node)
        d = [a.copy() for a in d6]
        d[2 * axis + 1][node] = 0.5 * h
        s = po.re_step(F, F < 0, dt, d, *SP, scheme=scheme)
        if moves:
            assert abs(s[node] - base[node]) > 1e-3 * dt[node], (offset, s[node], base[node])
        else:
            assert s[node] == base[node]
