"""vls_weno5_rk3_reinit_static (no device allocation; one static work array per
GPU) against the allocating path it stands for, vls_sdf3_create +
vls_sdf3_set_distance + vls_sdf3_weno5_rk3_reinit on a fresh handle, bit for
bit, over calls in a row on one GPU: the largest box the static memory holds,
then smaller ragged boxes (the array keeps the previous call's contents), and
its size check."""
import ctypes
import importlib

import numpy as np
import pytest
import torch

from golden import cases

vls = importlib.import_module("2018_7_vesicle_amd")

pytestmark = pytest.mark.gpu

STATIC_POINTS = 1 << 20  # VLS_STATIC_POINTS of include/vesicle_ls.h


def _box(rows, cols, pges, k):
    xv = np.linspace(-1.0, 1.0, cols)
    yv = np.linspace(-0.9 - 0.05 * k, 0.9, rows)
    zv = np.linspace(-1.1, 1.0 + 0.1 * k, pges)
    g = cases.Grid(xv, yv, zv)
    r = np.sqrt((g.X - 0.05 * k) ** 2 + g.Y ** 2 + (g.Z + 0.03 * k) ** 2)
    F = (r - 0.55) * (1 + 0.15 * np.sin(3 * g.X + k) * np.cos(2 * g.Y)) * (1 + 0.4 * g.X)
    return g, torch.from_numpy(np.ascontiguousarray(F)).to("cuda", torch.float64)


def _with_handle(lib, g, F, iterations, stream):
    h = ctypes.c_void_p()
    assert lib.vls_sdf3_create(ctypes.byref(h), g.rows, g.cols, g.pges, *g.spacing) == 0, lib.vls_last_error()
    try:
        assert lib.vls_sdf3_set_distance(h, F.data_ptr(), stream) == 0, lib.vls_last_error()
        assert lib.vls_sdf3_weno5_rk3_reinit(h, F.data_ptr(), iterations, stream) == 0, lib.vls_last_error()
        torch.cuda.synchronize()
    finally:
        lib.vls_sdf3_destroy(h)
    return F


@pytest.mark.parametrize("iterations", [12])
def test_static_reinit_equals_handle_reinit(iterations):
    lib = vls.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # the largest box, then smaller ones (rows not a multiple of the tile rows)
    for k, (rows, cols, pges) in enumerate([(128, 128, 64), (70, 66, 61), (33, 41, 29)]):
        assert rows * cols * pges <= STATIC_POINTS
        print(f"static reinit: box {rows}x{cols}x{pges}, {iterations} iterations", flush=True)
        g, F0 = _box(rows, cols, pges, k)
        Fs = F0.clone()
        rc = lib.vls_weno5_rk3_reinit_static(Fs.data_ptr(), iterations, rows, cols, pges, *g.spacing, stream)
        assert rc == 0, lib.vls_last_error()
        torch.cuda.synchronize()
        Fh = _with_handle(lib, g, F0.clone(), iterations, stream)
        assert not torch.equal(Fh, F0)
        assert torch.equal(Fs, Fh), f"box {rows}x{cols}x{pges}: max diff {float((Fs - Fh).abs().max())}"


def test_static_reinit_rejects_a_box_beyond_its_memory():
    lib = vls.lib()
    F = torch.ones(128 * 128 * 65, dtype=torch.float64, device="cuda")
    rc = lib.vls_weno5_rk3_reinit_static(F.data_ptr(), 2, 128, 128, 65, 0.1, 0.1, 0.1, None)
    assert rc == -1 and b"static work memory" in lib.vls_last_error()
    assert bool((F == 1).all())
