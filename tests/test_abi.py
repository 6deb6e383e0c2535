"""C-ABI checks that need no GPU: the library loads, exports every function
include/vesicle_ls.h declares (and the binding declares the same set), and
argument errors come back as VLS_ERR_INVALID before any device work."""
import ctypes
import importlib
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vesicle_ls.h")

vls = importlib.import_module("2018_7_vesicle_amd")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vls_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_api():
    names = declared()
    assert "vls_re_step" in names and "vls_sdf3_create" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = vls.lib()
    for name in declared():
        assert hasattr(lib, name), f"{name} not exported"


def test_binding_matches_header():
    assert sorted(vls.exported_names()) == declared()


def test_version_string():
    assert b"gfx950" in vls.lib().vls_version()


def test_invalid_geometry_rejected_without_device_work():
    lib = vls.lib()
    h = ctypes.c_void_p()
    rc = lib.vls_sdf3_create(ctypes.byref(h), 0, 8, 8, 0.1, 0.1, 0.1)
    assert rc == -1 and h.value is None
    assert b"empty" in lib.vls_last_error()
    rc = lib.vls_sdf3_create(ctypes.byref(h), 8, 8, 8, 0.1, -0.1, 0.1)
    assert rc == -1
    # num_ele mismatch on a drop-in entry point
    p = ctypes.c_void_p(16)
    rc = lib.vls_boundary_location(p, p, p, p, p, p, p, 10, 4, 4, 4, 0.1, 0.1, 0.1, None)
    assert rc == -1 and b"num_ele" in lib.vls_last_error()
    # null pointers
    z = ctypes.c_void_p()
    rc = lib.vls_advection_step(z, p, p, p, p, 4, 4, 4, 0.1, 0.1, 0.1, None)
    assert rc == -1 and b"null" in lib.vls_last_error()


def test_relaxation_entries_reject_bad_arguments_without_device_work():
    lib = vls.lib()
    p, z = ctypes.c_void_p(16), ctypes.c_void_p()
    # smoothers: null field, bad kind, non-finite dt, aliasing of the operator
    assert lib.vls_smooth_gmres(z, p, 4, 4, 4, 0.1, 0.1, 0.1, 1e-3, 0.5, 0, None) == -1
    assert b"null" in lib.vls_last_error()
    assert lib.vls_smooth_fft(p, p, 4, 4, 4, 0.1, 0.1, 0.1, 1e-3, 0.5, 2, None) == -1
    assert b"diffusion" in lib.vls_last_error()
    assert lib.vls_smooth_fft(p, p, 4, 4, 4, 0.1, 0.1, 0.1, float("nan"), 0.5, 0, None) == -1
    assert lib.vls_smooth_operator(p, p, 4, 4, 4, 0.1, 0.1, 0.1, 1e-3, 0.5, 0, None) == -1
    assert b"in place" in lib.vls_last_error()
    assert lib.vls_smooth_gmres(p, p, 0, 4, 4, 0.1, 0.1, 0.1, 1e-3, 0.5, 0, None) == -1
    # bending speed may not overwrite H
    assert lib.vls_bending_speed(p, p, ctypes.c_void_p(32), 4, 4, 4, 0.1, 0.1, 0.1, 1.0, 0.0, None) == -1
    # tension/pressure needs a handle; toolbox entries need every field
    assert lib.vls_sdf3_tension_pressure(z, p, p, p, p, p, p, 1.0, 1.0, 0.0, 0.5, p, p, None) == -1
    out = (ctypes.c_void_p * 30)(*([16] * 29 + [0]))
    fin = (ctypes.c_void_p * 14)(*([16] * 14))
    assert lib.vls_aux_calculus_toolbox(out, p, fin, 4, 4, 4, 0.1, 0.1, 0.1, 0.1, None) == -1
    assert b"null output 29" in lib.vls_last_error()


def test_batch_relax_rejects_bad_arguments_before_device_work():
    import torch

    with pytest.raises(vls.VLSError, match="one level set per grid"):
        vls.relax.BatchRelax([], [])
    g = vls.GD3([0.0, 0.1, 0.2, 0.3], [0.0, 0.1, 0.2, 0.3], [0.0, 0.1, 0.2, 0.3], device="cpu")
    with pytest.raises(vls.VLSError, match="one level set per grid"):
        vls.relax.BatchRelax([g, g], [torch.zeros(4, 4, 4, dtype=torch.float64)])
    with pytest.raises(vls.VLSError, match="FP64"):
        vls.relax.BatchRelax([g], [torch.zeros(4, 4, 4, dtype=torch.float32)])


def test_missing_library_fails_loudly(monkeypatch):
    lib_mod = importlib.import_module("2018_7_vesicle_amd._lib")
    monkeypatch.setattr(lib_mod, "_lib", None)
    monkeypatch.setattr(lib_mod, "LIB_PATH", "/nonexistent/libvesicle_ls.so")
    with pytest.raises(lib_mod.VLSError):
        lib_mod.lib()


def test_reinit_fp64_tail_setter_rejects_without_device_work():
    lib = vls.lib()
    assert lib.vls_sdf3_set_reinit_fp64_tail(ctypes.c_void_p(), 5) == -1
    assert b"null handle" in lib.vls_last_error()


def test_static_reinit_rejects_a_box_beyond_its_memory_without_device_work():
    lib = vls.lib()
    p = ctypes.c_void_p(16)
    # VLS_STATIC_POINTS = 2^20 points; 128 x 128 x 65 is one page more
    assert lib.vls_weno5_rk3_reinit_static(p, 2, 128, 128, 65, 0.1, 0.1, 0.1, None) == -1
    assert b"static work memory" in lib.vls_last_error()
    assert lib.vls_weno5_rk3_reinit_static(ctypes.c_void_p(), 2, 8, 8, 8, 0.1, 0.1, 0.1, None) == -1
    assert lib.vls_weno5_rk3_reinit_static(p, -1, 8, 8, 8, 0.1, 0.1, 0.1, None) == -1
