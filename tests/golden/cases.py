"""Input builders of the parity tests.

Fields are numpy float64 arrays of shape (pges, cols, rows) = (Nz, Nx, Ny), the
reference's column-major [rows, cols, pges] layout (MATLAB meshgrid gives
X(i,j,k) = xv(j), Y(i,j,k) = yv(i), Z(i,j,k) = zv(k); +SD/@GD3/GD3.m:86-105).
"""
import numpy as np


class Grid:
    """Uniform box grid; spacings as +SD/@GD3/GD3.m:90-93 computes them."""

    def __init__(self, xv, yv, zv):
        self.xv, self.yv, self.zv = (np.asarray(v, dtype=np.float64) for v in (xv, yv, zv))
        self.rows, self.cols, self.pges = len(self.yv), len(self.xv), len(self.zv)
        self.Z, self.X, self.Y = np.meshgrid(self.zv, self.xv, self.yv, indexing="ij")
        self.dx = self.xv[1] - self.xv[0]
        self.dy = self.yv[1] - self.yv[0]
        self.dz = self.zv[1] - self.zv[0]
        self.ds = (self.dx + self.dy + self.dz) / 3

    @property
    def shape(self):
        return (self.pges, self.cols, self.rows)

    @property
    def spacing(self):
        return (self.dx, self.dy, self.dz)


def linspace_grid(n, lo=-1.0, hi=1.0):
    v = np.linspace(lo, hi, n)
    return Grid(v, v, v)


def star_phi0(g, r0=0.6, amp=0.3):
    """(|x|-r0)(1+amp*sin(3 theta)cos(2 varphi))exp(0.5 x): theta = polar angle
    from +z, varphi = azimuth in the xy-plane."""
    r = np.sqrt(g.X ** 2 + g.Y ** 2 + g.Z ** 2)
    theta = np.arctan2(np.sqrt(g.X ** 2 + g.Y ** 2), g.Z)
    varphi = np.arctan2(g.Y, g.X)
    return (r - r0) * (1 + amp * np.sin(3 * theta) * np.cos(2 * varphi)) * np.exp(0.5 * g.X)


def merging_spheres(g, r0=0.6, zc=0.3):
    """min of two sphere distances (src/7_0_paper/1_0_WENO/p19_test_merging_front.m:6-67):
    has a kink on the plane z=0 inside the union."""
    f1 = np.sqrt(g.X ** 2 + g.Y ** 2 + (g.Z - zc) ** 2) - r0
    f2 = np.sqrt(g.X ** 2 + g.Y ** 2 + (g.Z + zc) ** 2) - r0
    return np.minimum(f1, f2)


def smooth_scalar(g):
    return 0.7 * g.Z + 0.3 * g.X * g.Y + 0.1 + 0.2 * np.sin(2.0 * g.Y)


def smooth_velocity(g):
    u = 0.5 * np.sin(1.3 * g.Y + 0.2) + 0.3 * g.Z
    v = -0.4 * np.cos(0.9 * g.X) + 0.25 * g.Z * g.X
    w = 0.6 * np.sin(1.1 * g.X - 0.7 * g.Y) - 0.2
    return [u, v, w]


def small_grids():
    """Small grids with distinct spacings per axis, so that a transposed axis
    or a swapped spacing cannot pass: a perturbed star, two merging spheres
    with a kink, and a field with exact zeros on a node plane."""
    g1 = Grid(np.linspace(-1.0, 1.0, 13), np.linspace(-0.9, 0.9, 12), np.linspace(-1.1, 1.1, 11))
    g2 = Grid(np.linspace(-1.0, 1.0, 12), np.linspace(-1.0, 1.0, 12), np.linspace(-1.25, 1.25, 14))
    g3 = Grid(np.linspace(-1.0, 1.0, 11), np.linspace(-0.8, 0.8, 10), np.linspace(-1.0, 1.0, 9))
    cases = {
        "star": (g1, star_phi0(g1)),
        "kink": (g2, merging_spheres(g2)),
        # exact zeros on the node plane j=6 plus a sign change elsewhere
        "plane": (g3, (g3.X - g3.xv[6]) * (1.0 + 0.5 * g3.Y ** 2) + 0.0 * g3.Z),
    }
    return cases


def matlab_sign(a):
    return np.sign(a)


def extension_velocity_from_normal(F, n3, floor):
    """+SD/@SDF3/WENO5RK3Extend.m:23-31 given upwind_normal output."""
    g = np.maximum(np.sqrt(n3[0] ** 2 + n3[1] ** 2 + n3[2] ** 2), floor)
    s = matlab_sign(F)
    return [s * (n3[0] / g), s * (n3[1] / g), s * (n3[2] / g)]
