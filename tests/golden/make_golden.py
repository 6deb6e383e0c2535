"""Write the golden vectors from the reference's own kernels.

Runs on a machine with an MI355X and oracle/_ref built (`make -C oracle ref` in a checkout
that has the reference sources).  For each case of cases.golden_grids() it
feeds seeded/analytic inputs to the reference kernels (compiled unmodified
from the reference's CUDA sources), in the order the reference's MATLAB
wrappers call them, and stores inputs and outputs as one .npz per case:

    python tests/golden/make_golden.py OUTDIR

The committed fixtures under tests/golden/ were produced by this script; the
CPU oracle is checked against them by tests/test_oracle_golden.py.  The script
also prints the oracle-vs-reference differences it sees.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)

import cases  # noqa: E402


def main(outdir):
    import torch

    from oracle import pyoracle as po
    from oracle.ref_kernels import RefKernels

    os.makedirs(outdir, exist_ok=True)
    ref = RefKernels()
    dev = torch.device("cuda:0")

    def T(a, dtype=torch.float64):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype=dtype)

    def N(t):
        return t.detach().cpu().numpy()

    for name, (g, F) in cases.golden_grids().items():
        print(f"case {name}: rows={g.rows} cols={g.cols} pges={g.pges}", flush=True)
        rows, cols, pges = g.rows, g.cols, g.pges
        dx, dy, dz = g.spacing
        rec = {"xv": g.xv, "yv": g.yv, "zv": g.zv, "F": F}
        Ft = T(F)

        # setDistance with the sixth-order locator and with ENO2 (+SD/@SDF3/setDistance.m:5-15)
        for key, fn in (("six", ref.boundary_location), ("eno", ref.boundary_correction)):
            d6 = [T(a) for a in po.prefilled_distances(F.shape, dx, dy, dz)]
            fn(d6, Ft, rows, cols, pges, dx, dy, dz)
            rec[f"d6_{key}"] = np.stack([N(t) for t in d6])
        md = rec["d6_six"].min(axis=0)
        md_eno = rec["d6_eno"].min(axis=0)
        mask = (F < 0).astype(np.uint8)
        deltat = 0.3 * md
        rec["mask"] = mask
        rec["deltat"] = deltat
        rec["deltat_eno"] = 0.3 * md_eno

        # re_step (WENO5) on F and on a perturbed stage field
        d6t = [T(a) for a in rec["d6_six"]]
        step = T(np.zeros_like(F))
        ref.re_step(step, Ft, T(mask, torch.uint8), T(deltat), d6t, rows, cols, pges, dx, dy, dz)
        rec["re_step"] = N(step)
        Fs = F + 0.3 * g.dx * np.sin(3.0 * g.X + 2.0 * g.Y - g.Z)
        rec["F_stage"] = Fs
        step = T(np.zeros_like(F))
        ref.re_step(step, T(Fs), T(mask, torch.uint8), T(deltat), d6t, rows, cols, pges, dx, dy, dz)
        rec["re_step_stage"] = N(step)

        # ENO2 re_step
        step = T(np.zeros_like(F))
        ref.eno2_re_step(step, Ft, T(mask, torch.uint8), T(rec["deltat_eno"]), [T(a) for a in rec["d6_eno"]],
                         rows, cols, pges, dx, dy, dz)
        rec["eno2_re_step"] = N(step)

        # extension pieces
        n3 = [T(np.zeros_like(F)) for _ in range(3)]
        ref.upwind_normal(n3, Ft, d6t, rows, cols, pges, dx, dy, dz)
        rec["upwind_normal"] = np.stack([N(t) for t in n3])
        C = cases.smooth_scalar(g)
        rec["C"] = C
        cp6 = [T(np.zeros_like(F)) for _ in range(6)]
        ref.boundary_interpolant(cp6, d6t, T(C), rows, cols, pges, dx, dy, dz)
        rec["cp6"] = np.stack([N(t) for t in cp6])
        v3 = cases.extension_velocity_from_normal(F, rec["upwind_normal"], 1e-6)
        rec["v3"] = np.stack(v3)
        step = T(np.zeros_like(F))
        ref.extend_step(step, T(deltat), T(C), [T(a) for a in v3], d6t, cp6, rows, cols, pges, dx, dy, dz)
        rec["extend_step"] = N(step)

        # advection operator and its twin
        u3 = cases.smooth_velocity(g)
        rec["u3"] = np.stack(u3)
        ast = T(np.zeros_like(F))
        ref.advection_step(ast, [T(a) for a in u3], Ft, rows, cols, pges, dx, dy, dz)
        rec["advection_step"] = N(ast)
        wd = [T(np.zeros_like(F)) for _ in range(6)]
        ref.weno_derivative(wd, Ft, rows, cols, pges, dx, dy, dz)
        rec["weno_derivative"] = np.stack([N(t) for t in wd])

        # calculus toolbox with HPrimal = max(F,0) (+SD/@SDF3/GPUsetCalculusToolBox.m:4)
        tb = [T(np.zeros_like(F)) for _ in range(18)]
        ref.set_calculus_toolbox(tb, Ft, T(np.maximum(F, 0)), rows, cols, pges, dx, dy, dz, g.ds)
        rec["toolbox"] = np.stack([N(t) for t in tb])
        torch.cuda.synchronize()

        np.savez_compressed(os.path.join(outdir, f"golden_{name}.npz"), **rec)

        # oracle vs reference, printed for the record
        d6o = po.prefilled_distances(F.shape, dx, dy, dz)
        po.boundary_location(d6o, F, dx, dy, dz)
        print(f"  boundary_location max|diff| {np.abs(np.stack(d6o) - rec['d6_six']).max():.3e}")
        so = po.re_step(F, mask, deltat, list(rec["d6_six"]), dx, dy, dz)
        print(f"  re_step max|diff| {np.abs(so - rec['re_step']).max():.3e} "
              f"(max|ref| {np.abs(rec['re_step']).max():.3e})")
        eo = po.extend_step(deltat, C, list(rec["v3"]), list(rec["d6_six"]), list(rec["cp6"]), dx, dy, dz)
        print(f"  extend_step max|diff| {np.abs(eo - rec['extend_step']).max():.3e}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "gpurun_out/golden")
