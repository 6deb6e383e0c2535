"""The benchmark configurations' own outputs, on the GPU, against the oracle.

Config 1: 64^3 on [-1,1]^3, phi0 = (|x|-0.6)(1+0.3 sin3theta cos2varphi) e^{0.5x},
ENO2 with the subcell fix, forward Euler, CFL 0.5, 100 iterations; reports
max and mean of ||grad phi|-1| in |phi| < 3dx and the interface displacement
from the sphere r = 0.6 (the zero set of phi0).

Config 3 (small): the vesicle-step pipeline in FP32 bulk against the same
pipeline in FP64 from one seed; the error bounds are stated in the test.
"""
import importlib

import numpy as np
import pytest
import torch

import cases
from oracle import pyoracle as po

vls = importlib.import_module("2018_7_vesicle_amd")
diag = importlib.import_module("2018_7_vesicle_amd.diagnostics")
pytestmark = pytest.mark.gpu


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", torch.float64)


def test_config1_eno2_euler_reinit():
    g = cases.linspace_grid(64)
    F0 = np.ascontiguousarray(cases.star_phi0(g))
    G = vls.GD3(g.xv, g.yv, g.zv)
    m = vls.SDF3(G, T(F0))
    Fg = m.ENORK2Reinitialization(T(F0), 100, rk2=False, cfl=0.5)
    torch.cuda.synchronize()
    ref = po.eno2_reinit(F0, 100, *g.spacing, rk2=False, cfl=0.5)
    assert np.array_equal(Fg.cpu().numpy(), ref)
    e_max0, e_mean0, _ = diag.grad_error(G, T(F0))
    e_max, e_mean, nband = diag.grad_error(G, Fg)
    disp, ncross = diag.sphere_displacement(G, Fg, 0.6)
    disp0, _ = diag.sphere_displacement(G, T(F0), 0.6)
    print(f"config1: ||grad|-1| max {e_max:.3e} mean {e_mean:.3e} (phi0: {e_max0:.3e} {e_mean0:.3e}); "
          f"displacement {disp / g.dx:.3e} dx over {ncross} crossings (phi0: {disp0 / g.dx:.3e} dx)")
    assert e_mean < 0.1 * e_mean0
    assert disp < 0.05 * g.dx


def test_config3_precision_comparison_small():
    """The config-3 comparison at 64^3 (8 modes, 10 steps, reinit every 5):
    the FP32-bulk run stays within 1e-4 dx of the FP64 run normal to the
    interface at the 99th percentile, area/volume within 1e-6 relative, the
    same crossings exist in both, and a one-rounding perturbation of the
    initial data moves the FP64 interface by less than the FP32 run does."""
    diag = importlib.import_module("2018_7_vesicle_amd.diagnostics")
    r = diag.precision_comparison(64, 8, 10, conditioning=True, initial_reinit=20)
    e = r["errors_fp32_vs_fp64"]
    d = e["interface_displacement"]
    assert d["unmatched"] == 0
    assert d["normal"]["p99_dx"] < 1e-4, d
    assert d["normal"]["max_dx"] <= d["along_line"]["max_dx"] * (1 + 1e-12)
    assert e["area_rel_drift_max"] < 1e-6 and e["volume_rel_drift_max"] < 1e-6, e
    c = r["fp64_rounded_start_vs_fp64"]["interface_displacement"]
    assert c["normal"]["max_dx"] <= d["normal"]["max_dx"]
