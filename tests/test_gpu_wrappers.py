"""Parity of the @SDF3 wrapper entry points (C-ABI handle) with the oracle's
restatement of the MATLAB wrappers, iteration loops included.

Bar: bit-for-bit equality over the whole box after the iteration loops (the
kernels and the oracle evaluate the same IEEE sequence).  The area and volume
sums are compared at 1e-13 relative: summation order differs and the cosine
delta/H calls cos/sin, whose last bit differs between the device library and
glibc.  Bit-exactness matters here: the reference's quintic interface
interpolant is ill-conditioned in monomial form, so a 1-ulp difference in its
input moves its output by ~1e-11 relative; a merely close implementation of
the surface redistance cannot agree with the oracle to better than that.
"""
import importlib

import numpy as np
import pytest
import torch

import cases
from oracle import pyoracle as po

vls = importlib.import_module("2018_7_vesicle_amd")
pytestmark = pytest.mark.gpu


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", torch.float64)


def N(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def close(a, b, rtol=1e-12, what=""):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max()
    assert err <= rtol * scale, f"{what}: max|diff| {err:.3e} > {rtol:.0e} * {scale:.3e}"


def exact(a, b, what=""):
    """bit-for-bit equality (the kernels and the oracle evaluate the same IEEE sequence)"""
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        n = int((a != b).sum())
        scale = max(np.abs(b).max(), 1e-300)
        raise AssertionError(f"{what}: {n} of {a.size} values differ, max|diff| {d.max():.3e} "
                             f"({d.max() / scale:.1e} relative)")


def same_sign(a, b, what=""):
    bad = np.sign(a) != np.sign(b)
    assert not bad.any(), f"{what}: {bad.sum()} sign flips, smallest |value| {np.abs(b[bad]).min():.3e}"


def make(grid_kind):
    if grid_kind == "star":
        g = cases.Grid(np.linspace(-1.0, 1.0, 36), np.linspace(-0.95, 0.95, 32), np.linspace(-1.05, 1.05, 30))
        F = cases.star_phi0(g)
    else:
        g = cases.Grid(np.linspace(-1.0, 1.0, 32), np.linspace(-1.0, 1.0, 34), np.linspace(-1.25, 1.25, 40))
        F = cases.merging_spheres(g)
    G = vls.GD3(g.xv, g.yv, g.zv)
    return g, np.ascontiguousarray(F), G


@pytest.fixture(params=["star", "kink"])
def setup(request):
    return make(request.param)


def test_set_distance(setup):
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    m.setDistance()
    got = [N(t) for t in m.distances()]
    d6, md = po.set_distance(F, *g.spacing)
    for k in range(6):
        exact(got[k], d6[k], f"distance {k}")
    exact(got[6], md, "min_dist")


def test_weno5_rk3_reinitialization(setup):
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    m.setDistance()
    got = N(m.WENO5RK3Reinitialization(T(F), 6))
    d6, md = po.set_distance(F, *g.spacing)
    ref = po.weno5_rk3_reinit(F, 6, d6, md, *g.spacing)
    exact(got, ref, "WENO5RK3Reinitialization")


@pytest.mark.parametrize("rk2,cfl,iters", [(False, 0.5, 12), (True, 0.3, 6)])
def test_eno2_reinitialization(setup, rk2, cfl, iters):
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    got = N(m.ENORK2Reinitialization(T(F), iters, rk2=rk2, cfl=cfl))
    ref = po.eno2_reinit(F, iters, *g.spacing, rk2=rk2, cfl=cfl)
    exact(got, ref, "ENORK2Reinitialization")


def test_weno5_rk3_extend(setup):
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    m.setDistance()
    C = cases.smooth_scalar(g)
    got = N(m.WENO5RK3Extend(T(C), 6))
    d6, md = po.set_distance(F, *g.spacing)
    ref = po.weno5_rk3_extend(C, 6, F, d6, md, *g.spacing)
    exact(got, ref, "WENO5RK3Extend")


def test_wenork3_reinitialization(setup):
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    got = N(m.WENORK3Reinitialization(T(F), 6))
    ref = po.wenork3_reinit(F, 6, *g.spacing)
    exact(got, ref, "WENORK3Reinitialization")


@pytest.mark.parametrize("scheme", ["enork2", "wenork3"])
def test_variant_extend(setup, scheme):
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    C = cases.smooth_scalar(g)
    fn = m.ENORK2Extend if scheme == "enork2" else m.WENORK3Extend
    got = N(fn(T(C), 6))
    ref = po.variant_extend(C, 6, F, *g.spacing, scheme)
    exact(got, ref, f"{scheme} extension")


def test_surface_redistance(setup):
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    A = g.X - 0.2 * np.cos(2 * np.pi * g.Y)
    got = N(m.WENO5RK3ClosetPointSurfaceRedistance(T(A), 2, 3))
    ref = po.surface_redistance(A, F, 2, 3, *g.spacing)
    exact(got, ref, "surface redistance")


def test_advection_rk3(setup):
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    u3 = cases.smooth_velocity(g)
    dt = 0.4 * g.dx / max(np.abs(u).max() for u in u3)
    got = N(m.WENO5RK3Advection(T(F), *[T(u) for u in u3], dt, 4))
    ref = po.advect_rk3(F, 4, dt, u3, *g.spacing)
    exact(got, ref, "advection")


def test_area_volume(setup):
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    a, v = m.calAreaVolume()
    ra, rv = po.area_volume_towers(F, *g.spacing)
    assert abs(a - ra) <= 1e-13 * abs(ra) and abs(v - rv) <= 1e-13 * abs(rv)
    eps = 1.5 * g.dx
    a, v = m.cosineAreaVolume(eps)
    ra, rv = po.area_volume_cosine(F, *g.spacing, eps)
    assert abs(a - ra) <= 1e-13 * abs(ra) and abs(v - rv) <= 1e-13 * abs(rv)


def test_calculus_toolbox_object(setup):
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    m.GPUsetCalculusToolBox()
    ref = po.set_calculus_toolbox(F, np.maximum(F, 0), *g.spacing, g.ds)
    for name in ("Heaviside", "DiracDelta", "MeanCurvature", "FGradMag"):
        exact(N(getattr(m, name)), ref[name], name)


def test_aux_calculus_toolbox_object(setup):
    """GPUAsetCalculusToolBox.m as a whole: AHPrimal, the kernel's 27 fields and
    the periodic-difference tail (AFDiracDelta, ADiracDeltaDn)"""
    g, F, G = setup
    m = vls.SDF3(G, T(F))
    with pytest.raises(vls.VLSError):
        m.GPUAsetCalculusToolBox()
    A = np.ascontiguousarray(0.6 * g.Z + 0.25 * g.X * g.Y - 0.15 * g.Y + 0.05)
    m.A = T(A)
    with pytest.raises(vls.VLSError):
        m.GPUAsetCalculusToolBox()
    m.GPUsetCalculusToolBox()
    m.GPUAsetCalculusToolBox()
    ref = po.aux_calculus_toolbox(A, po.set_calculus_toolbox(F, np.maximum(F, 0), *g.spacing, g.ds), *g.spacing,
                                  g.ds)
    assert len(ref) == 30
    for name, want in ref.items():
        exact(N(getattr(m, name)), want, name)


def test_pipeline_small():
    """two vesicle steps of the benchmark pipeline (one with reinit and surface
    redistance) at 32^3 against the oracle's restatement of the same pipeline"""
    cfg = vls.pipeline.PipelineConfig(n=32, modes=4, ext_iters=4, reinit_every=2, reinit_iters=4, redist_iters=(2, 2),
                                      initial_reinit=4)
    p = vls.pipeline.Pipeline(cfg)
    p.run(2)
    xv, F0, psi0, u = vls.pipeline.make_inputs(cfg)
    dx = xv[1] - xv[0]
    F, psi, Vn, av = po.pipeline(F0, psi0, u, p.dt, p.eps, (dx, dx, dx), 2, cfg.ext_iters, cfg.reinit_every,
                                 cfg.reinit_iters, cfg.redist_iters, cfg.initial_reinit)
    exact(N(p.map.F), F, "pipeline F")
    exact(N(p.psi), psi, "pipeline psi")
    exact(N(p.Vn), Vn, "pipeline Vn")
    got_av = N(p.area_volume)
    assert np.all(np.abs(got_av - av) <= 1e-13 * np.abs(av)), (got_av, av)


@pytest.mark.parametrize("shape", [(24, 20, 2), (30, 26, 4), (9, 5, 11), (70, 3, 8)])
def test_thin_and_small_boxes(shape):
    """z-marching edge cases: fewer pages than the z-window (clamped at both
    ends at once), a box smaller than one tile, partial tiles in rows and
    columns; reinitialization, extension and advection against the oracle"""
    rows, cols, pges = shape
    g = cases.Grid(np.linspace(-1.0, 1.0, cols), np.linspace(-0.9, 1.0, rows), np.linspace(-1.1, 0.9, pges))
    G = vls.GD3(g.xv, g.yv, g.zv)
    F = np.ascontiguousarray(np.sqrt(g.X ** 2 + g.Y ** 2 + (0.3 * g.Z) ** 2) - 0.55)
    m = vls.SDF3(G, T(F))
    m.setDistance()
    got = N(m.WENO5RK3Reinitialization(T(F), 3))
    d6, md = po.set_distance(F, *g.spacing)
    exact(got, po.weno5_rk3_reinit(F, 3, d6, md, *g.spacing), f"reinit {shape}")
    C = cases.smooth_scalar(g)
    got = N(m.WENO5RK3Extend(T(C), 3))
    exact(got, po.weno5_rk3_extend(C, 3, F, d6, md, *g.spacing), f"extend {shape}")
    u3 = cases.smooth_velocity(g)
    dt = 0.4 * g.dx / max(np.abs(u).max() for u in u3)
    got = N(m.WENO5RK3Advection(T(F), *[T(u) for u in u3], dt, 2))
    exact(got, po.advect_rk3(F, 2, dt, u3, *g.spacing), f"advection {shape}")


def test_field_without_interface():
    """no sign change anywhere: an empty interface list in the fused launches"""
    g = cases.Grid(np.linspace(-1, 1, 34), np.linspace(-1, 1, 30), np.linspace(-1, 1, 26))
    G = vls.GD3(g.xv, g.yv, g.zv)
    F = np.ascontiguousarray(np.sqrt(g.X ** 2 + g.Y ** 2 + g.Z ** 2) + 0.2)
    m = vls.SDF3(G, T(F))
    m.setDistance()
    got = N(m.WENO5RK3Reinitialization(T(F), 4))
    d6, md = po.set_distance(F, *g.spacing)
    exact(got, po.weno5_rk3_reinit(F, 4, d6, md, *g.spacing), "reinit")
    C = cases.smooth_scalar(g)
    got = N(m.WENO5RK3Extend(T(C), 4))
    exact(got, po.weno5_rk3_extend(C, 4, F, d6, md, *g.spacing), "extend")
