"""The CPU oracle of the vesicle relaxation step's operators (@GD3 smoothers,
bending speed, tension/pressure solve) against what the mathematics fixes:
Fourier modes through the smoothers, residuals of the sparse operator built as
+SD/@GD3/GD3.m builds it, and the 2x2 solve against Cramer's rule."""
import math

import numpy as np
import pytest

import cases
from oracle import relax_oracle as ro

RNG = np.random.default_rng(20181207)


def _periodic_grid(rows, cols, pges, dx, dy, dz):
    # nodes 0..n-1 times h: the box of period n*h that the wavenumbers assume
    return cases.Grid(np.arange(cols) * dx, np.arange(rows) * dy, np.arange(pges) * dz)


def test_wavenumbers_even_sizes_are_the_dft_frequencies():
    for n, h in ((8, 0.1), (16, 0.037), (2, 1.0)):
        assert np.allclose(ro.gd3_wavenumbers(n, h), 2 * math.pi * np.fft.fftfreq(n, h), rtol=0, atol=1e-12)


def test_wavenumbers_odd_sizes_keep_the_half_integer_offsets():
    # -5/2 : 5/2-1 = [-2.5 -1.5 -0.5 0.5 1.5], fftshift moves the last two first
    k = ro.gd3_wavenumbers(5, 1.0) / (2 * math.pi / 5)
    assert np.allclose(k, [0.5, 1.5, -2.5, -1.5, -0.5], rtol=0, atol=1e-14)


@pytest.mark.parametrize("diffusion", [False, True])
def test_smooth_fft_scales_a_fourier_mode(diffusion):
    rows, cols, pges, dx, dy, dz = 12, 16, 10, 0.11, 0.07, 0.13
    g = _periodic_grid(rows, cols, pges, dx, dy, dz)
    mx, my, mz = 3, 2, 1
    kx, ky, kz = (2 * math.pi * mx / (cols * dx), 2 * math.pi * my / (rows * dy), 2 * math.pi * mz / (pges * dz))
    f = np.cos(kx * g.X + ky * g.Y - kz * g.Z + 0.3)
    dt, alpha = 2e-3, 0.5
    k2 = kx * kx + ky * ky + kz * kz
    want = f / (1 + alpha * dt * (k2 if diffusion else k2 * k2))
    got = ro.smooth_fft(f, dx, dy, dz, dt, alpha, diffusion)
    assert np.abs(got - want).max() < 1e-13


def test_laplacian_matrix_is_the_periodic_stencil():
    rows, cols, pges, dx, dy, dz = 6, 7, 5, 0.2, 0.3, 0.25
    f = RNG.normal(size=(pges, cols, rows))
    L = ro.gd3_laplacian_matrix(f.shape, dx, dy, dz)
    got = (L @ f.ravel()).reshape(f.shape)
    want = ro.gd3_laplacian(f, dx, dy, dz)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    # one interior point by hand: x neighbours are +-1 along cols (axis 1)
    p, c, r = 2, 3, 1
    hand = ((f[p, c + 1, r] - 2 * f[p, c, r] + f[p, c - 1, r]) / dx ** 2
            + (f[p, c, r + 1] - 2 * f[p, c, r] + f[p, c, r - 1]) / dy ** 2
            + (f[p + 1, c, r] - 2 * f[p, c, r] + f[p - 1, c, r]) / dz ** 2)
    assert abs(want[p, c, r] - hand) < 1e-12 * abs(hand)


@pytest.mark.parametrize("diffusion", [False, True])
@pytest.mark.parametrize("dims", [(12, 10, 8), (9, 7, 6)])
def test_exact_solve_and_gmres_solve_the_sparse_system(diffusion, dims):
    rows, cols, pges = dims
    dx, dy, dz = 0.09, 0.11, 0.1
    s = RNG.normal(size=(pges, cols, rows))
    dt, alpha = 4e-4 if not diffusion else 5e-2, 0.5
    A = ro.smooth_matrix(s.shape, dx, dy, dz, dt, alpha, diffusion)
    x = ro.smooth_exact(s, dx, dy, dz, dt, alpha, diffusion)
    res = np.linalg.norm(s.ravel() - A @ x.ravel()) / np.linalg.norm(s)
    assert res < 1e-14
    xg, it, rr = ro.smooth_gmres(s, dx, dy, dz, dt, alpha, diffusion)
    assert rr <= 1e-12 and it < 300
    # gmres stops on the preconditioned residual: the true one is larger by
    # up to the preconditioner's norm (1 + alpha dt |k|^4 reaches ~300 here)
    assert np.linalg.norm(s.ravel() - A @ xg.ravel()) / np.linalg.norm(s) < 1e-10
    assert np.abs(xg - x).max() < 1e-11 * np.abs(x).max()
    # the spectral preconditioner alone is not the solution: GMRES had work to do
    xs = ro.smooth_fft(s, dx, dy, dz, dt, alpha, diffusion)
    assert np.abs(xs - x).max() > 1e-6 * np.abs(x).max()
    assert it > 2


def test_bending_speed_constants_and_a_mode():
    rows, cols, pges, dx, dy, dz = 8, 10, 6, 0.1, 0.12, 0.15
    g = _periodic_grid(rows, cols, pges, dx, dy, dz)
    H = np.full(g.shape, 1.7)
    G = np.full(g.shape, 0.4)
    v = ro.bending_speed(H, G, dx, dy, dz, 1.0, 0.3)
    want = 0.5 * 1.7 * (1.7 ** 2 - 4 * 0.4) + 2 * 0.3 * 0.4 - 0.5 * 0.09 * 1.7
    assert np.abs(v - want).max() < 1e-14
    kx = 2 * math.pi / (cols * dx)
    Hm = np.sin(kx * g.X)
    lap = ro.gd3_laplacian(Hm, dx, dy, dz)
    # the periodic second difference of a mode: (2cos(k dx) - 2)/dx^2 times the mode
    assert np.abs(lap - (2 * math.cos(kx * dx) - 2) / dx ** 2 * Hm).max() < 1e-12


def test_tension_pressure_solves_the_2x2_system():
    g = cases.linspace_grid(24)
    F = np.sqrt(g.X ** 2 + g.Y ** 2 + g.Z ** 2) - 0.6
    ds = g.ds
    band = np.abs(F) < 2 * ds
    dirac = np.where(band, np.cos(F / ds) ** 2 / ds, 0.0)
    heav = np.clip(0.5 + F / (4 * ds), 0, 1)
    gmag = 1 + 0.01 * g.X
    H = -2 / 0.6 + 0.1 * g.Z
    Vb = 0.3 * np.sin(2 * g.X) + 0.05
    speed, r = ro.tension_pressure(F, H, Vb, dirac, gmag, heav, ds, 4.4, 0.9, cfl=0.05)
    a = np.array([[r["c11"], r["c12"]], [r["c12"], 4.4]])
    det = a[0, 0] * a[1, 1] - a[0, 1] * a[1, 0]
    T = (r["s1"] * a[1, 1] - a[0, 1] * r["s2"]) / det
    P = (a[0, 0] * r["s2"] - a[1, 0] * r["s1"]) / det
    assert abs(r["Tension"] - T) < 1e-12 * abs(T) and abs(r["Pressure"] - P) < 1e-12 * abs(P)
    assert r["Dt"] == 0.05 * ds / np.abs(Vb[band]).max()
    assert np.abs(speed - (T * H - Vb + P) * gmag).max() < 1e-12 * np.abs(speed).max()


def test_relax_step_reproduces_the_reference_grid_study():
    """the oracle's composition of one relaxation step (48^3, CFL 0.1) against
    the first entry of the reference's own record of that run
    (tests/golden/oblate48.mat, p00_test_singlephase_relax.m:260-270): the
    oblate ellipsoid, setDistance and 500 reinitialization iterations, then
    one loop body.  Dt rests on max|Vb| over the band, which agrees to 2e-7."""
    import importlib
    import os

    vls = importlib.import_module("2018_7_vesicle_amd")
    from oracle import pyoracle as po

    rec = vls.matio.loadmat(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oblate48.mat"))
    xv, yv, zv, F0 = vls.relax.oblate_ellipsoid(48)
    dx, dy, dz = xv[1] - xv[0], yv[1] - yv[0], zv[1] - zv[0]
    d6, md = po.set_distance(F0, dx, dy, dz)
    F = po.weno5_rk3_reinit(F0, 500, d6, md, dx, dy, dz)
    a0, v0 = po.area_volume_towers(F, dx, dy, dz)
    _, info = ro.relax_step(F, a0, v0, dx, dy, dz, cfl=0.1)
    assert abs(info["Pressure"] - rec["P48"][0, 0]) <= 1e-9 * abs(rec["P48"][0, 0])
    assert abs(info["Tension"] - rec["ten48"][0, 0]) <= 1e-9 * abs(rec["ten48"][0, 0])
    assert abs(info["Dt"] - rec["t48"][0, 0]) <= 1e-6 * rec["t48"][0, 0]
