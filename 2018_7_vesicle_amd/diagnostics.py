"""Error measures of a level set, as the benchmark configurations report them.

* |grad phi| - 1 in a band |phi| < k*dx, with grad phi from the clamped
  central differences of set_calculus_toolbox (FGradMag,
  src/7_0_paper/2_0_Vesicle/CUDA_Code/4_0_Calculus_ToolBox/set_calculus_toolbox.cu:123-127);
* interface crossings: where the zero set cuts grid lines, located by
  setDistance's sixth-order subcell locator (boundary_location.cu:84-132),
  as points in space;
* the displacement of those crossings from a reference (an exact sphere, or
  the crossings of another field on the same grid).

Everything heavy runs through the HIP library; the host only gathers the
crossing list.
"""
import numpy as np
import torch

from .sdf3 import GD3, SDF3, TOOLBOX_FIELDS, feval_set_calculus_toolbox


def grad_error(grid: GD3, F: torch.Tensor, band_dx: float = 3.0):
    """max and mean of ||grad F| - 1| over |F| < band_dx * Dx"""
    outs = [grid.zeros() for _ in TOOLBOX_FIELDS]
    hp = torch.clamp(F, min=0.0).contiguous()
    feval_set_calculus_toolbox(outs, F, hp, grid.mrows, grid.ncols, grid.lshts, grid.Dx, grid.Dy, grid.Dz, grid.Ds)
    g = outs[3]
    band = F.abs() < band_dx * grid.Dx
    e = (g[band] - 1.0).abs()
    return float(e.max()), float(e.mean()), int(band.sum())


def crossing_points(grid: GD3, F: torch.Tensor):
    """(points[n,3], keys[n,2]) of the grid-line crossings of the zero set of F.

    key = (axis, linear index of the node on the low side); points in (x, y, z)."""
    m = SDF3(grid, F)
    m.setDistance()
    xpr, _, ypf, _, zpu, _, _ = m.distances()
    torch.cuda.synchronize()
    X, Y, Z = (t.cpu().numpy() for t in grid.meshgrid())
    pts, keys = [], []
    for axis, (d, sp) in enumerate(((xpr, grid.Dx), (ypf, grid.Dy), (zpu, grid.Dz))):
        dn = d.cpu().numpy()
        idx = np.flatnonzero(dn.ravel() < sp)
        p = np.stack([X.ravel()[idx], Y.ravel()[idx], Z.ravel()[idx]], axis=1)
        p[:, axis] += dn.ravel()[idx]
        pts.append(p)
        keys.append(np.stack([np.full(idx.shape, axis), idx], axis=1))
    del m
    return np.concatenate(pts), np.concatenate(keys)


def sphere_displacement(grid: GD3, F: torch.Tensor, radius: float, center=(0.0, 0.0, 0.0)):
    """max over crossings of the distance to the sphere, | |p - c| - R |"""
    pts, _ = crossing_points(grid, F)
    r = np.linalg.norm(pts - np.asarray(center)[None, :], axis=1)
    return float(np.abs(r - radius).max()), len(pts)


def crossing_displacement(grid: GD3, Fa: torch.Tensor, Fb: torch.Tensor):
    """max |p_a - p_b| over the grid-line crossings both fields have, and the
    number of crossings only one of them has"""
    pa, ka = crossing_points(grid, Fa)
    pb, kb = crossing_points(grid, Fb)
    ia = {(int(a), int(i)): n for n, (a, i) in enumerate(ka)}
    common = [(ia[(int(a), int(i))], n) for n, (a, i) in enumerate(kb) if (int(a), int(i)) in ia]
    if not common:
        return float("nan"), len(pa) + len(pb)
    sa, sb = zip(*common)
    d = np.linalg.norm(pa[list(sa)] - pb[list(sb)], axis=1)
    return float(d.max()), len(pa) + len(pb) - 2 * len(common)
