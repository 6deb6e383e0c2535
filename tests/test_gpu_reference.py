"""The HIP product against the reference's own kernels, live, at sizes beyond
the committed fixtures.

oracle/_ref holds the reference's CUDA sources compiled unmodified for gfx950
(`make -C oracle ref`, only where the reference sources exist); the tests are
skipped when it was not built.  Same bar as against the oracle: 1e-12 of the
field's largest magnitude, distances to 1e-12*ds, masks exact.
"""
import importlib

import numpy as np
import pytest
import torch

import cases
from oracle import pyoracle as po
from oracle import ref_kernels

vls = importlib.import_module("2018_7_vesicle_amd")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref_kernels.available(), reason="oracle/_ref (reference kernels) not built")]


def T(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def N(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def close(a, b, rtol=1e-12, what=""):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max()
    assert err <= rtol * scale, f"{what}: max|diff| {err:.3e} > {rtol:.0e} * {scale:.3e}"


@pytest.fixture(scope="module")
def ref():
    return ref_kernels.RefKernels()


@pytest.fixture(scope="module", params=["config1_64", "kink_aniso"])
def case(request):
    if request.param == "config1_64":
        g = cases.linspace_grid(64)
        F = cases.star_phi0(g)
    else:
        g = cases.Grid(np.linspace(-1.0, 1.0, 56), np.linspace(-1.0, 1.0, 60), np.linspace(-1.3, 1.3, 64))
        F = cases.merging_spheres(g)
    return g, np.ascontiguousarray(F)


def test_distances_and_re_step(ref, case):
    g, F = case
    dx, dy, dz = g.spacing
    r, c, p = g.rows, g.cols, g.pges
    d_ref = [T(a) for a in po.prefilled_distances(F.shape, dx, dy, dz)]
    ref.boundary_location(d_ref, T(F), r, c, p, dx, dy, dz)
    d_got = [T(a) for a in po.prefilled_distances(F.shape, dx, dy, dz)]
    vls.feval_boundary_location(*d_got, T(F), F.size, r, c, p, dx, dy, dz)
    dr, dg = [N(t) for t in d_ref], [N(t) for t in d_got]
    for k in range(6):
        s = (dx, dx, dy, dy, dz, dz)[k]
        assert np.array_equal(dg[k] < s, dr[k] < s)
        assert np.abs(dg[k] - dr[k]).max() <= 1e-12 * s
    md = np.min(np.stack(dr), axis=0)
    mask = T((F < 0).astype(np.uint8), torch.uint8)
    dt = T(0.3 * md)
    s_ref = torch.zeros_like(T(F))
    ref.re_step(s_ref, T(F), mask, dt, d_ref, r, c, p, dx, dy, dz)
    s_got = torch.zeros_like(T(F))
    vls.feval_re_step(s_got, T(F), mask, dt, *d_ref, r, c, p, dx, dy, dz, F.size)
    a, b = N(s_got), N(s_ref)
    assert np.array_equal(a == 0, b == 0)
    close(a, b, what="re_step vs reference")


def test_extension_kernels(ref, case):
    g, F = case
    dx, dy, dz = g.spacing
    r, c, p = g.rows, g.cols, g.pges
    d6 = [T(a) for a in po.prefilled_distances(F.shape, dx, dy, dz)]
    ref.boundary_location(d6, T(F), r, c, p, dx, dy, dz)
    n_ref = [torch.zeros_like(T(F)) for _ in range(3)]
    ref.upwind_normal(n_ref, T(F), d6, r, c, p, dx, dy, dz)
    n_got = [torch.zeros_like(T(F)) for _ in range(3)]
    vls.feval_upwind_normal(*n_got, T(F), *d6, r, c, p, dx, dy, dz, F.size)
    for a, b in zip(n_got, n_ref):
        close(N(a), N(b), what="upwind_normal vs reference")
    C = T(cases.smooth_scalar(g))
    cp_ref = [torch.zeros_like(C) for _ in range(6)]
    ref.boundary_interpolant(cp_ref, d6, C, r, c, p, dx, dy, dz)
    cp_got = [torch.zeros_like(C) for _ in range(6)]
    vls.feval_boundary_interpolant(*cp_got, *d6, C, r, c, p, dx, dy, dz, F.size)
    for a, b in zip(cp_got, cp_ref):
        close(N(a), N(b), rtol=1e-11, what="boundary_interpolant vs reference")
    v3 = [T(a) for a in cases.extension_velocity_from_normal(F, [N(t) for t in n_ref], 1e-6)]
    dt = T(0.3 * np.min(np.stack([N(t) for t in d6]), axis=0))
    s_ref = torch.zeros_like(C)
    ref.extend_step(s_ref, dt, C, v3, d6, cp_ref, r, c, p, dx, dy, dz)
    s_got = torch.zeros_like(C)
    vls.feval_extend_step(s_got, dt, C, *v3, *d6, *cp_ref, r, c, p, dx, dy, dz, F.size)
    close(N(s_got), N(s_ref), what="extend_step vs reference")


def test_variant_kernels(ref, case):
    """WENO-RK3 / ENO-RK2 variants against weno_rk3_reinitialization.cu,
    enork2_extend.cu and wenork3_extend.cu"""
    g, F = case
    dx, dy, dz = g.spacing
    r, c, p = g.rows, g.cols, g.pges
    d_ref = [T(a) for a in po.prefilled_distances(F.shape, dx, dy, dz)]
    ref.cubic_boundary_correction(d_ref, T(F), r, c, p, dx, dy, dz)
    d_got = [T(a) for a in po.prefilled_distances(F.shape, dx, dy, dz)]
    vls.feval_cubic_boundary_correction(*d_got, T(F), F.size, r, c, p, dx, dy, dz)
    dr, dg = [N(t) for t in d_ref], [N(t) for t in d_got]
    for k in range(6):
        s = (dx, dx, dy, dy, dz, dz)[k]
        assert np.array_equal(dg[k] < s, dr[k] < s)
        assert np.abs(dg[k] - dr[k]).max() <= 1e-12 * s
    mask = T((F < 0).astype(np.uint8), torch.uint8)
    dt = T(0.3 * np.min(np.stack(dr), axis=0))
    s_ref = torch.zeros_like(T(F))
    ref.wenork3_re_step(s_ref, T(F), mask, dt, d_ref, r, c, p, dx, dy, dz)
    s_got = torch.zeros_like(T(F))
    vls.feval_wenork3_re_step(s_got, T(F), mask, dt, *d_ref, r, c, p, dx, dy, dz, F.size)
    a, b = N(s_got), N(s_ref)
    assert np.array_equal(a == 0, b == 0)
    close(a, b, what="wenork3 re_step vs reference")
    C = T(cases.smooth_scalar(g))
    for scheme in ("enork2", "wenork3"):
        d6 = d_ref
        if scheme == "enork2":
            d6 = [T(a) for a in po.prefilled_distances(F.shape, dx, dy, dz)]
            ref.boundary_correction(d6, T(F), r, c, p, dx, dy, dz)
        n_ref = [torch.zeros_like(T(F)) for _ in range(3)]
        ref.variant_upwind_normal(scheme, n_ref, T(F), d6, r, c, p, dx, dy, dz)
        n_got = [torch.zeros_like(T(F)) for _ in range(3)]
        getattr(vls, f"feval_{scheme}_upwind_normal")(*n_got, T(F), *d6, r, c, p, dx, dy, dz, F.size)
        for x, y in zip(n_got, n_ref):
            close(N(x), N(y), what=f"{scheme} upwind_normal vs reference")
        cp_ref = [torch.zeros_like(C) for _ in range(6)]
        ref.variant_boundary_interpolate(scheme, cp_ref, d6, C, r, c, p, dx, dy, dz)
        cp_got = [torch.zeros_like(C) for _ in range(6)]
        getattr(vls, f"feval_{scheme}_boundary_interpolate")(*cp_got, *d6, C, r, c, p, dx, dy, dz, F.size)
        for x, y in zip(cp_got, cp_ref):
            close(N(x), N(y), rtol=1e-11, what=f"{scheme} boundary_interpolate vs reference")
        v3 = [T(a) for a in cases.extension_velocity_from_normal(F, [N(t) for t in n_ref], 1e-6)]
        dte = T(0.3 * np.min(np.stack([N(t) for t in d6]), axis=0))
        s_ref = torch.zeros_like(C)
        ref.variant_extend_step(scheme, s_ref, dte, C, v3, d6, cp_ref, r, c, p, dx, dy, dz)
        s_got = torch.zeros_like(C)
        getattr(vls, f"feval_{scheme}_extend_step")(s_got, dte, C, *v3, *d6, *cp_ref, r, c, p, dx, dy, dz, F.size)
        close(N(s_got), N(s_ref), what=f"{scheme} extend_step vs reference")


def test_advection_toolbox(ref, case):
    g, F = case
    dx, dy, dz = g.spacing
    r, c, p = g.rows, g.cols, g.pges
    u3 = [T(a) for a in cases.smooth_velocity(g)]
    a_ref = torch.zeros_like(T(F))
    ref.advection_step(a_ref, u3, T(F), r, c, p, dx, dy, dz)
    a_got = torch.zeros_like(T(F))
    vls.feval_advection_step(a_got, *u3, T(F), r, c, p, dx, dy, dz)
    close(N(a_got), N(a_ref), what="advection_step vs reference")
    hp = T(np.maximum(F, 0))
    t_ref = [torch.zeros_like(T(F)) for _ in range(18)]
    ref.set_calculus_toolbox(t_ref, T(F), hp, r, c, p, dx, dy, dz, g.ds)
    t_got = [torch.zeros_like(T(F)) for _ in range(18)]
    vls.feval_set_calculus_toolbox(t_got, T(F), hp, r, c, p, dx, dy, dz, g.ds)
    for name, a, b in zip(vls.TOOLBOX_FIELDS, t_got, t_ref):
        close(N(a), N(b), rtol=1e-11, what=f"{name} vs reference")


def test_aux_toolbox(ref, case):
    """auxi_set_calculus_toolbox.cu, fed the reference's own toolbox of F"""
    g, F = case
    dx, dy, dz = g.spacing
    r, c, p = g.rows, g.cols, g.pges
    A = np.ascontiguousarray(0.6 * g.Z + 0.25 * g.X * g.Y - 0.15 * g.Y + 0.05)
    ft = [torch.zeros_like(T(F)) for _ in range(18)]
    ref.set_calculus_toolbox(ft, T(F), T(np.maximum(F, 0)), r, c, p, dx, dy, dz, g.ds)
    ahp = T(np.maximum(A, 0))
    a_ref = [torch.zeros_like(T(F)) for _ in range(27)]
    ref.auxi_set_calculus_toolbox(a_ref, T(A), ahp, ft[:13], r, c, p, dx, dy, dz, g.ds)
    a_got = [torch.zeros_like(T(F)) for _ in range(27)]
    vls.feval_auxi_set_calculus_toolbox(a_got, T(A), ahp, ft[:13], r, c, p, dx, dy, dz, g.ds)
    for name, a, b in zip(vls.AUX_TOOLBOX_FIELDS, a_got, a_ref):
        close(N(a), N(b), rtol=1e-11, what=f"{name} vs reference")
