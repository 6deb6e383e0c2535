"""Error measures of a level set, as the benchmark configurations report them.

* |grad phi| - 1 in a band |phi| < k*dx, with grad phi from the clamped
  central differences of set_calculus_toolbox (FGradMag,
  src/7_0_paper/2_0_Vesicle/CUDA_Code/4_0_Calculus_ToolBox/set_calculus_toolbox.cu:31-35);
* interface crossings: where the zero set cuts grid lines, located by
  setDistance's sixth-order subcell locator (boundary_location.cu:84-132),
  as points in space;
* the displacement of those crossings from a reference (an exact sphere, or
  the crossings of another field on the same grid).

Everything heavy runs through the HIP library; the host only gathers the
crossing list.

precision_comparison runs the vesicle-step pipeline twice from the same seed
and initial state, FP64 throughout and FP32 bulk, and states the FP32 run's
errors against the FP64 run.
"""
import time

import numpy as np
import torch

from .sdf3 import GD3, SDF3, TOOLBOX_FIELDS, feval_set_calculus_toolbox


def grad_error(grid: GD3, F: torch.Tensor, band_dx: float = 3.0):
    """max and mean of ||grad F| - 1| over |F| < band_dx * Dx"""
    F = F.double()
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
    m = SDF3(grid, F.double().contiguous())
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


def crossing_displacements(grid: GD3, Fa: torch.Tensor, Fb: torch.Tensor, normal=False):
    """|p_a - p_b| at every grid-line crossing both fields have, and the number
    of crossings only one of them has.

    normal=True gives the component of p_a - p_b along the unit normal of Fb
    instead (central differences of Fb at the crossing's low node): both
    crossings lie on one grid line, so that is |p_a - p_b| |n_axis|, the
    distance of p_a from Fb's zero set to first order.  Along the line, a
    crossing where the grid line grazes the interface (n_axis -> 0) moves by
    (change of F)/|dF/daxis| however small the change; the normal component
    does not have that amplification."""
    pa, ka = crossing_points(grid, Fa)
    pb, kb = crossing_points(grid, Fb)
    n = grid.NumElt
    _, ia, ib = np.intersect1d(ka[:, 0] * n + ka[:, 1], kb[:, 0] * n + kb[:, 1], return_indices=True)
    d = np.linalg.norm(pa[ia] - pb[ib], axis=1)
    if normal:
        F = Fb.double()
        g = [torch.zeros_like(F) for _ in range(3)]
        sp = (grid.Dx, grid.Dy, grid.Dz)
        dims = (1, 2, 0)  # tensor dims of x (cols), y (rows), z (pges)
        for a in range(3):
            dm = dims[a]
            hi = torch.cat([F.narrow(dm, 1, F.shape[dm] - 1), F.narrow(dm, F.shape[dm] - 1, 1)], dim=dm)
            lo = torch.cat([F.narrow(dm, 0, 1), F.narrow(dm, 0, F.shape[dm] - 1)], dim=dm)
            g[a] = (hi - lo) / (2 * sp[a])
        idx = torch.from_numpy(kb[ib, 1]).to(F.device)
        axis = torch.from_numpy(kb[ib, 0]).to(F.device)
        gx, gy, gz = (t.reshape(-1)[idx] for t in g)
        gm = torch.sqrt(gx * gx + gy * gy + gz * gz).clamp_min(1e-14)
        ga = torch.where(axis == 0, gx, torch.where(axis == 1, gy, gz)).abs()
        d = d * (ga / gm).cpu().numpy()
    return d, len(pa) + len(pb) - 2 * len(ia)


def crossing_displacement(grid: GD3, Fa: torch.Tensor, Fb: torch.Tensor):
    """max |p_a - p_b| over the grid-line crossings both fields have, and the
    number of crossings only one of them has"""
    d, unmatched = crossing_displacements(grid, Fa, Fb)
    return (float(d.max()) if len(d) else float("nan")), unmatched


def _displacement_stats(grid, Fa, Fb):
    """displacement of Fa's interface from Fb's, in units of dx: along the grid
    lines and normal to the interface (see crossing_displacements)"""
    out = {}
    for name, normal in (("along_line", False), ("normal", True)):
        d, unmatched = crossing_displacements(grid, Fa, Fb, normal=normal)
        d = d / grid.Dx
        out[name] = {"max_dx": float(d.max()), "p99_dx": float(np.percentile(d, 99)), "mean_dx": float(d.mean()),
                     "count_above_1e-4_dx": int((d > 1e-4).sum())}
    out["crossings"] = int(len(d))
    out["unmatched"] = int(unmatched)
    return out


def precision_comparison(n=256, modes=32, steps=200, conditioning=False, **cfg_kw):
    """The vesicle-step pipeline (pipeline.Pipeline) at n^3 with a `modes`-mode
    velocity, run for `steps` steps in FP64 and in FP32 bulk from the same
    seed and initial state; the FP64 run is the reference (the FP64 path
    reproduces the CPU oracle bit for bit).  Returns the throughput of each run
    and the FP32 run's errors: interface displacement (over the grid-line
    crossings both fields have, in units of dx: max, 99th percentile, mean),
    |grad phi| - 1 in the band |phi| < 3 dx, and the relative area/volume
    drift (max over the steps).

    conditioning=True adds a third run, FP64 started from the initial state
    rounded to FP32, and its displacement from the FP64 run: how far the
    problem itself carries a one-rounding perturbation of the data, the floor
    under any FP32 result."""
    from .pipeline import Pipeline, PipelineConfig

    out = {"grid": n, "modes": modes, "steps": steps}
    runs = {}
    precs = ("fp64", "fp32", "fp64_rounded_start") if conditioning else ("fp64", "fp32")
    for prec in precs:
        cfg = PipelineConfig(n=n, modes=modes, precision="fp64" if prec == "fp64_rounded_start" else prec,
                             **cfg_kw)
        p = Pipeline(cfg)
        if prec == "fp64_rounded_start":
            p.map.F = p.map.F.float().double()
            p.psi = p.psi.float().double()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        updates = p.run(steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        runs[prec] = p
        out[prec] = {"wall_s": wall, "ms_per_step": wall / steps * 1e3, "updates_per_s": updates / wall}
    out["fp32_speedup"] = out["fp64"]["wall_s"] / out["fp32"]["wall_s"]
    p64, p32 = runs["fp64"], runs["fp32"]
    grid = p64.grid
    F64, F32 = p64.map.F, p32.map.F.double()
    g64, g32 = grad_error(grid, F64), grad_error(grid, F32)
    av64, av32 = p64.area_volume.cpu().numpy(), p32.area_volume.cpu().numpy()
    rel = np.abs(av32 - av64) / np.abs(av64)
    band = (F64.abs() < 3 * grid.Dx)
    out["errors_fp32_vs_fp64"] = {
        "interface_displacement": _displacement_stats(grid, F32, F64),
        "phi_max_abs_diff_band_dx": float((F32 - F64).abs()[band].max()) / grid.Dx,
        "grad_err_max": [g64[0], g32[0]],
        "grad_err_mean": [g64[1], g32[1]],
        "area_rel_drift_max": float(rel[:, 0].max()),
        "volume_rel_drift_max": float(rel[:, 1].max()),
        "area_rel_drift_final": float(rel[-1, 0]),
        "volume_rel_drift_final": float(rel[-1, 1]),
    }
    out["area_volume_final_fp64"] = [float(av64[-1, 0]), float(av64[-1, 1])]
    if conditioning:
        pr = runs["fp64_rounded_start"]
        avr = pr.area_volume.cpu().numpy()
        relr = np.abs(avr - av64) / np.abs(av64)
        out["fp64_rounded_start_vs_fp64"] = {
            "interface_displacement": _displacement_stats(grid, pr.map.F, F64),
            "area_rel_drift_max": float(relr[:, 0].max()),
            "volume_rel_drift_max": float(relr[:, 1].max()),
        }
    return out
