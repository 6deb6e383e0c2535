"""Write the golden vectors of the auxiliary-field toolbox from the reference's
own kernels.

Runs on a machine with an MI355X and oracle/_ref built (`make -C oracle ref` in a checkout
that has the reference sources; the code objects are set_calculus_toolbox.co
and auxi_set_calculus_toolbox.co).  For each case of cases.golden_grids() it
runs set_calculus_toolbox on F and then auxi_set_calculus_toolbox on the
auxiliary field A with the toolbox of F, in the order
+SD/@SDF3/GPUsetCalculusToolBox.m and GPUAsetCalculusToolBox.m:40-54 call them,
and stores inputs and outputs as one aux_<case>.npz:

    python tests/golden/make_golden_aux.py OUTDIR

The committed fixtures tests/golden/aux_*.npz were produced by this script;
tests/test_oracle_golden.py checks the CPU oracle against them.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)

import cases  # noqa: E402


def aux_field(g):
    """an auxiliary level set that cuts the surfaces of every golden case"""
    return 0.6 * g.Z + 0.25 * g.X * g.Y - 0.15 * g.Y + 0.05


def main(outdir):
    import torch

    from oracle import pyoracle as po
    from oracle.ref_kernels import RefKernels

    os.makedirs(outdir, exist_ok=True)
    ref = RefKernels()
    dev = torch.device("cuda:0")

    def T(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype=torch.float64)

    def N(t):
        return t.detach().cpu().numpy()

    for name, (g, F) in cases.golden_grids().items():
        print(f"case {name}: rows={g.rows} cols={g.cols} pges={g.pges}", flush=True)
        rows, cols, pges = g.rows, g.cols, g.pges
        dx, dy, dz = g.spacing
        A = aux_field(g)
        ft = [T(np.zeros_like(F)) for _ in range(18)]
        ref.set_calculus_toolbox(ft, T(F), T(np.maximum(F, 0)), rows, cols, pges, dx, dy, dz, g.ds)
        out = [T(np.zeros_like(F)) for _ in range(27)]
        ref.auxi_set_calculus_toolbox(out, T(A), T(np.maximum(A, 0)), ft[:13], rows, cols, pges, dx, dy, dz, g.ds)
        torch.cuda.synchronize()
        rec = {"xv": g.xv, "yv": g.yv, "zv": g.zv, "F": F, "A": A,
               "ftool": np.stack([N(t) for t in ft]), "aux": np.stack([N(t) for t in out])}
        np.savez_compressed(os.path.join(outdir, f"aux_{name}.npz"), **rec)

        # oracle vs reference, printed for the record
        o = po.auxi_set_calculus_toolbox(A, np.maximum(A, 0), list(rec["ftool"][:13]), dx, dy, dz, g.ds)
        for k, field in enumerate(po.AUX_TOOLBOX_FIELDS):
            d = np.abs(o[field] - rec["aux"][k]).max()
            s = max(np.abs(rec["aux"][k]).max(), 1e-300)
            print(f"  {field:18s} max|diff| {d:.3e} rel {d / s:.1e}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else HERE)
