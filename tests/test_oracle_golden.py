"""The CPU oracle against the golden vectors written by the reference's own
kernels (tests/golden/make_golden.py ran them, compiled unmodified from the
reference's CUDA sources, on an MI355X).

Tolerances: the reference evaluates the same formulas in a different order
(FMA contraction, pow()), so agreement is to rounding.  Distances are pinned to
1e-12*ds absolute (the monomial-form quintic root moves by ~1e-13*ds under
rounding); every other field to 1e-12 of the field's largest magnitude.  Masks
(which distances are below ds, the frozen set) must agree exactly.
"""
import glob
import os

import numpy as np
import pytest

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "golden_*.npz")))


def _load(path):
    z = np.load(path)  # allow_pickle=False by default
    return {k: z[k] for k in z.files}


def _spacing(rec):
    return (rec["xv"][1] - rec["xv"][0], rec["yv"][1] - rec["yv"][0], rec["zv"][1] - rec["zv"][0])


def _close(a, b, rtol=1e-12):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max()
    assert err <= rtol * scale, f"max|diff| {err:.3e} > {rtol:.0e} * {scale:.3e}"


@pytest.fixture(scope="module", params=FILES, ids=[os.path.basename(f)[7:-4] for f in FILES])
def rec(request):
    return _load(request.param)


def test_fixture_count():
    assert len(FILES) == 3


def test_boundary_location(rec):
    F = rec["F"]
    dx, dy, dz = _spacing(rec)
    d6 = po.prefilled_distances(F.shape, dx, dy, dz)
    po.boundary_location(d6, F, dx, dy, dz)
    ds = np.array([dx, dx, dy, dy, dz, dz])[:, None, None, None]
    ref = rec["d6_six"]
    got = np.stack(d6)
    assert np.array_equal(got < ds, ref < ds)
    assert np.abs(got - ref).max() <= 1e-12 * min(dx, dy, dz)


def test_boundary_correction_eno2(rec):
    F = rec["F"]
    dx, dy, dz = _spacing(rec)
    d6 = po.prefilled_distances(F.shape, dx, dy, dz)
    po.boundary_location(d6, F, dx, dy, dz, order=2)
    ref = rec["d6_eno"]
    assert np.abs(np.stack(d6) - ref).max() <= 1e-12 * min(dx, dy, dz)


def test_re_step_weno5(rec):
    dx, dy, dz = _spacing(rec)
    for lsf, key in ((rec["F"], "re_step"), (rec["F_stage"], "re_step_stage")):
        got = po.re_step(lsf, rec["mask"], rec["deltat"], list(rec["d6_six"]), dx, dy, dz)
        assert np.array_equal(got == 0, rec[key] == 0)
        _close(got, rec[key])


def test_re_step_eno2(rec):
    dx, dy, dz = _spacing(rec)
    got = po.re_step(rec["F"], rec["mask"], rec["deltat_eno"], list(rec["d6_eno"]), dx, dy, dz, scheme="eno2")
    _close(got, rec["eno2_re_step"])


def test_upwind_normal(rec):
    dx, dy, dz = _spacing(rec)
    got = po.upwind_normal(rec["F"], list(rec["d6_six"]), dx, dy, dz)
    for a, b in zip(got, rec["upwind_normal"]):
        _close(a, b)


def test_boundary_interpolant(rec):
    dx, dy, dz = _spacing(rec)
    got = po.boundary_interpolant(list(rec["d6_six"]), rec["C"], dx, dy, dz)
    for a, b in zip(got, rec["cp6"]):
        assert np.array_equal(a == 0, b == 0)
        _close(a, b, rtol=1e-11)


def test_extend_step(rec):
    dx, dy, dz = _spacing(rec)
    got = po.extend_step(rec["deltat"], rec["C"], list(rec["v3"]), list(rec["d6_six"]), list(rec["cp6"]),
                         dx, dy, dz)
    _close(got, rec["extend_step"])


def test_advection_step(rec):
    dx, dy, dz = _spacing(rec)
    got = po.advection_step(list(rec["u3"]), rec["F"], dx, dy, dz)
    _close(got, rec["advection_step"])


def test_weno_derivative(rec):
    dx, dy, dz = _spacing(rec)
    got = po.weno_derivative(rec["F"], dx, dy, dz)
    for a, b in zip(got, rec["weno_derivative"]):
        _close(a, b)


def test_set_calculus_toolbox(rec):
    dx, dy, dz = _spacing(rec)
    ds = (dx + dy + dz) / 3
    F = rec["F"]
    got = po.set_calculus_toolbox(F, np.maximum(F, 0), dx, dy, dz, ds)
    for k, name in enumerate(po.TOOLBOX_FIELDS):
        _close(got[name], rec["toolbox"][k], rtol=1e-11)


# --------------------------------------------------------------------------
# WENO-RK3 / ENO-RK2 variants: vectors written by the reference's
# weno_rk3_reinitialization.cu, enork2_extend.cu and wenork3_extend.cu
# (tests/golden/make_golden_variants.py)
# --------------------------------------------------------------------------
VFILES = sorted(glob.glob(os.path.join(HERE, "golden", "variants_*.npz")))


@pytest.fixture(scope="module", params=VFILES, ids=[os.path.basename(f)[9:-4] for f in VFILES])
def vrec(request):
    return _load(request.param)


def test_variant_fixture_count():
    assert len(VFILES) == 3


def test_cubic_distances(vrec):
    dx, dy, dz = _spacing(vrec)
    d6 = po.prefilled_distances(vrec["F"].shape, dx, dy, dz)
    po.boundary_location(d6, vrec["F"], dx, dy, dz, order=3)
    for k in range(6):
        s = (dx, dx, dy, dy, dz, dz)[k]
        assert np.array_equal(d6[k] < s, vrec["d6_cubic"][k] < s)
        assert np.abs(d6[k] - vrec["d6_cubic"][k]).max() <= 1e-12 * s


def test_wenork3_re_step(vrec):
    dx, dy, dz = _spacing(vrec)
    for lsf, key in ((vrec["F"], "wenork3_re_step"), (vrec["F_stage"], "wenork3_re_step_stage")):
        got = po.re_step(lsf, vrec["mask"], vrec["deltat_cubic"], list(vrec["d6_cubic"]), dx, dy, dz,
                         scheme="wenork3")
        assert np.array_equal(got == 0, vrec[key] == 0)
        _close(got, vrec[key])


@pytest.mark.parametrize("scheme,dkey,tkey", [("enork2", "d6_eno", "deltat_eno"),
                                              ("wenork3", "d6_cubic", "deltat_cubic")])
def test_variant_extension_kernels(vrec, scheme, dkey, tkey):
    dx, dy, dz = _spacing(vrec)
    d6 = list(vrec[dkey])
    got = po.variant_upwind_normal(vrec["F"], d6, dx, dy, dz, scheme)
    for a, b in zip(got, vrec[f"{scheme}_upwind_normal"]):
        _close(a, b)
    cp = po.variant_boundary_interpolate(d6, vrec["C"], dx, dy, dz, scheme)
    for a, b in zip(cp, vrec[f"{scheme}_cp6"]):
        assert np.array_equal(a == 0, b == 0)
        _close(a, b, rtol=1e-11)
    st = po.variant_extend_step(vrec[tkey], vrec["C"], list(vrec[f"{scheme}_v3"]), d6, list(vrec[f"{scheme}_cp6"]),
                                dx, dy, dz, scheme)
    _close(st, vrec[f"{scheme}_extend_step"])


# --------------------------------------------------------------------------
# auxiliary-field toolbox: vectors written by the reference's
# set_calculus_toolbox.cu and auxi_set_calculus_toolbox.cu
# (tests/golden/make_golden_aux.py); the oracle is fed the reference's own
# toolbox of F
# --------------------------------------------------------------------------
AFILES = sorted(glob.glob(os.path.join(HERE, "golden", "aux_*.npz")))


def test_aux_fixture_count():
    assert len(AFILES) == 3


@pytest.mark.parametrize("path", AFILES, ids=[os.path.basename(f)[4:-4] for f in AFILES])
def test_auxi_set_calculus_toolbox(path):
    arec = _load(path)
    dx, dy, dz = _spacing(arec)
    ds = (dx + dy + dz) / 3
    A = arec["A"]
    got = po.auxi_set_calculus_toolbox(A, np.maximum(A, 0), list(arec["ftool"][:13]), dx, dy, dz, ds)
    for k, name in enumerate(po.AUX_TOOLBOX_FIELDS):
        _close(got[name], arec["aux"][k], rtol=1e-11)
    # and the oracle's own toolbox of F is the reference's
    ft = po.set_calculus_toolbox(arec["F"], np.maximum(arec["F"], 0), dx, dy, dz, ds)
    for k, name in enumerate(po.TOOLBOX_FIELDS):
        _close(ft[name], arec["ftool"][k], rtol=1e-11)
