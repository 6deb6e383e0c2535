"""Write the golden vectors of the WENO-RK3 / ENO-RK2 variants from the
reference's own kernels.

Runs on a machine with an MI355X and oracle/_ref built (`make -C oracle ref` in a checkout
that has the reference sources; the variant code objects are
weno_rk3_reinitialization.co, enork2_extend.co and wenork3_extend.co).  For
each case of cases.golden_grids() it runs the reference kernels in the order
+SD/@SDF3/WENORK3Reinitialization.m, ENORK2Extend.m and WENORK3Extend.m call
them and stores inputs and outputs as one variants_<case>.npz:

    python tests/golden/make_golden_variants.py OUTDIR

The committed fixtures tests/golden/variants_*.npz were produced by this
script; tests/test_oracle_golden.py checks the CPU oracle against them.
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
        Ft = T(F)
        C = cases.smooth_scalar(g)
        Fs = F + 0.3 * g.dx * np.sin(3.0 * g.X + 2.0 * g.Y - g.Z)
        rec = {"xv": g.xv, "yv": g.yv, "zv": g.zv, "F": F, "C": C, "F_stage": Fs}
        mask = (F < 0).astype(np.uint8)
        rec["mask"] = mask

        # distances: cubic (weno_rk3_reinitialization.cu boundary_correction) and
        # quadratic ENO (1_0 boundary_correction, the ENORK2 extension's)
        d6 = [T(a) for a in po.prefilled_distances(F.shape, dx, dy, dz)]
        ref.cubic_boundary_correction(d6, Ft, rows, cols, pges, dx, dy, dz)
        rec["d6_cubic"] = np.stack([N(t) for t in d6])
        d6e = [T(a) for a in po.prefilled_distances(F.shape, dx, dy, dz)]
        ref.boundary_correction(d6e, Ft, rows, cols, pges, dx, dy, dz)
        rec["d6_eno"] = np.stack([N(t) for t in d6e])
        rec["deltat_cubic"] = 0.3 * rec["d6_cubic"].min(axis=0)
        rec["deltat_eno"] = 0.3 * rec["d6_eno"].min(axis=0)

        # WENO-RK3 re_step on F and on a perturbed stage field
        for key, lsf in (("wenork3_re_step", F), ("wenork3_re_step_stage", Fs)):
            step = T(np.zeros_like(F))
            ref.wenork3_re_step(step, T(lsf), T(mask, torch.uint8), T(rec["deltat_cubic"]), d6, rows, cols, pges,
                                dx, dy, dz)
            rec[key] = N(step)

        # the two extensions: upwind normal of F, interface values of C, extend_step
        for scheme, dk, dtk, dd in (("enork2", "d6_eno", "deltat_eno", d6e),
                                    ("wenork3", "d6_cubic", "deltat_cubic", d6)):
            n3 = [T(np.zeros_like(F)) for _ in range(3)]
            ref.variant_upwind_normal(scheme, n3, Ft, dd, rows, cols, pges, dx, dy, dz)
            rec[f"{scheme}_upwind_normal"] = np.stack([N(t) for t in n3])
            cp6 = [T(np.zeros_like(F)) for _ in range(6)]
            ref.variant_boundary_interpolate(scheme, cp6, dd, T(C), rows, cols, pges, dx, dy, dz)
            rec[f"{scheme}_cp6"] = np.stack([N(t) for t in cp6])
            v3 = cases.extension_velocity_from_normal(F, rec[f"{scheme}_upwind_normal"], 1e-6)
            rec[f"{scheme}_v3"] = np.stack(v3)
            step = T(np.zeros_like(F))
            ref.variant_extend_step(scheme, step, T(rec[dtk]), T(C), [T(a) for a in v3], dd, cp6, rows, cols, pges,
                                    dx, dy, dz)
            rec[f"{scheme}_extend_step"] = N(step)
        torch.cuda.synchronize()

        np.savez_compressed(os.path.join(outdir, f"variants_{name}.npz"), **rec)

        # oracle vs reference, printed for the record
        d6o = po.prefilled_distances(F.shape, dx, dy, dz)
        po.boundary_location(d6o, F, dx, dy, dz, order=3)
        print(f"  cubic distances max|diff| {np.abs(np.stack(d6o) - rec['d6_cubic']).max():.3e}")
        so = po.re_step(F, mask, rec["deltat_cubic"], list(rec["d6_cubic"]), dx, dy, dz, scheme="wenork3")
        print(f"  wenork3 re_step max|diff| {np.abs(so - rec['wenork3_re_step']).max():.3e} "
              f"(max|ref| {np.abs(rec['wenork3_re_step']).max():.3e})")
        for scheme, dk, dtk in (("enork2", "d6_eno", "deltat_eno"), ("wenork3", "d6_cubic", "deltat_cubic")):
            st = po.variant_extend_step(rec[dtk], C, list(rec[f"{scheme}_v3"]), list(rec[dk]),
                                        list(rec[f"{scheme}_cp6"]), dx, dy, dz, scheme)
            print(f"  {scheme} extend_step max|diff| {np.abs(st - rec[f'{scheme}_extend_step']).max():.3e} "
                  f"(max|ref| {np.abs(rec[f'{scheme}_extend_step']).max():.3e})", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else HERE)
