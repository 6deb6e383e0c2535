"""ctypes binding of libvesicle_ls.so (the C ABI declared in include/vesicle_ls.h).

The library is built in-tree by build.py.  Loading it never falls back to
anything else: if it is missing, every entry point raises.  torch is imported
first so that the process has one HIP runtime (torch's), which the library
then shares; torch is used only for device memory and streams.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library links against)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvesicle_ls.so")

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
PD = ctypes.POINTER(ctypes.c_double)

# name -> argument types (the trailing P of most entries is the hipStream_t)
SIGNATURES = {
    "vls_last_error": [],
    "vls_version": [],
    "vls_boundary_location": [P] * 7 + [I] * 4 + [D] * 3 + [P],
    "vls_boundary_correction": [P] * 7 + [I] * 4 + [D] * 3 + [P],
    "vls_re_step": [P] * 10 + [I] * 3 + [D] * 3 + [I, P],
    "vls_eno2_re_step": [P] * 10 + [I] * 3 + [D] * 3 + [I, P],
    "vls_upwind_normal": [P] * 10 + [I] * 3 + [D] * 3 + [I, P],
    "vls_boundary_interpolant": [P] * 13 + [I] * 3 + [D] * 3 + [I, P],
    "vls_extend_step": [P] * 18 + [I] * 3 + [D] * 3 + [I, P],
    "vls_advection_step": [P] * 5 + [I] * 3 + [D] * 3 + [P],
    "vls_weno_derivative": [P] * 7 + [I] * 3 + [D] * 3 + [P],
    "vls_set_calculus_toolbox": [P] * 3 + [I] * 3 + [D] * 4 + [I, P],
    "vls_auxi_set_calculus_toolbox": [P] * 4 + [I] * 3 + [D] * 4 + [I, P],
    "vls_aux_calculus_toolbox": [P] * 3 + [I] * 3 + [D] * 4 + [P],
    "vls_mat_count": [ctypes.c_char_p, P],
    "vls_mat_info": [ctypes.c_char_p, I, ctypes.c_char_p, I, P, P, P],
    "vls_mat_read": [ctypes.c_char_p, ctypes.c_char_p, P, ctypes.c_int64],
    "vls_mat_write": [ctypes.c_char_p, I, P, P, P, P, I],
    "vls_smooth_fft": [P] * 2 + [I] * 3 + [D] * 5 + [I, P],
    "vls_smooth_gmres": [P] * 2 + [I] * 3 + [D] * 5 + [I, P],
    "vls_smooth_operator": [P] * 2 + [I] * 3 + [D] * 5 + [I, P],
    "vls_bending_speed": [P] * 3 + [I] * 3 + [D] * 5 + [P],
    "vls_sdf3_tension_pressure": [P] * 7 + [D] * 4 + [P] * 3,
    "vls_cubic_boundary_correction": [P] * 7 + [I] * 4 + [D] * 3 + [P],
    "vls_wenork3_re_step": [P] * 10 + [I] * 3 + [D] * 3 + [I, P],
    "vls_enork2_upwind_normal": [P] * 10 + [I] * 3 + [D] * 3 + [I, P],
    "vls_enork2_boundary_interpolate": [P] * 13 + [I] * 3 + [D] * 3 + [I, P],
    "vls_enork2_extend_step": [P] * 18 + [I] * 3 + [D] * 3 + [I, P],
    "vls_wenork3_upwind_normal": [P] * 10 + [I] * 3 + [D] * 3 + [I, P],
    "vls_wenork3_boundary_interpolate": [P] * 13 + [I] * 3 + [D] * 3 + [I, P],
    "vls_wenork3_extend_step": [P] * 18 + [I] * 3 + [D] * 3 + [I, P],
    "vls_sdf3_create": [P, I, I, I, D, D, D],
    "vls_sdf3_destroy": [P],
    "vls_sdf3_set_distance": [P, P, P],
    "vls_sdf3_distances": [P, P],
    "vls_sdf3_weno5_rk3_reinit": [P, P, I, P],
    "vls_weno5_rk3_reinit_static": [P, I, I, I, I, D, D, D, P],
    "vls_sdf3_eno2_reinit": [P, P, I, I, D, P],
    "vls_sdf3_wenork3_reinit": [P, P, I, P],
    "vls_sdf3_enork2_extend": [P, P, P, I, P],
    "vls_sdf3_wenork3_extend": [P, P, P, I, P],
    "vls_sdf3_weno5_rk3_extend": [P, P, P, I, P],
    "vls_sdf3_surface_redistance": [P, P, P, I, I, P],
    "vls_sdf3_advect_rk3": [P, P, P, P, P, D, I, P],
    "vls_sdf3_max_speed": [P, P, P, P, PD, P],
    "vls_sdf3_area_volume_towers": [P, P, PD, PD, P],
    "vls_sdf3_area_volume_cosine": [P, P, D, PD, PD, P],
    "vls_sdf3_area_volume_device": [P, P, I, D, P, P],
    "vls_sdf3_normal_velocity": [P] * 7,
    "vls_sdf3_set_distance_f32": [P, P, P],
    "vls_sdf3_weno5_rk3_reinit_f32": [P, P, I, P],
    "vls_sdf3_weno5_rk3_extend_f32": [P, P, P, I, P],
    "vls_sdf3_surface_redistance_f32": [P, P, P, I, I, P],
    "vls_sdf3_advect_rk3_f32": [P, P, P, P, P, D, I, P],
    "vls_sdf3_max_speed_f32": [P, P, P, P, PD, P],
    "vls_sdf3_area_volume_device_f32": [P, P, I, D, P, P],
    "vls_sdf3_normal_velocity_f32": [P] * 7,
    "vls_set_graphs": [I],
    "vls_probe_valu_rate": [I, PD, P],
    "vls_probe_stream_rate": [ctypes.c_int64, PD, P],
    "vls_batch_create": [P, I, P, P, P, P, P, P],
    "vls_batch_destroy": [P],
    "vls_batch_layout": [P, P, P],
    "vls_batch_set_blocks": [P, I],
    "vls_batch_get_blocks": [P, P],
    "vls_batch_set_tile_rows": [P, I],
    "vls_batch_get_tile_rows": [P, P],
    "vls_batch_set_zmarch_depth": [P, I],
    "vls_batch_get_zmarch_depth": [P, P],
    "vls_sdf3_set_tile_rows": [P, I],
    "vls_sdf3_set_reinit_fp64_tail": [P, I],
    "vls_sdf3_set_zmarch_depth": [P, I],
    "vls_sdf3_get_launch": [P, P, P],
    "vls_batch_autotune": [P, P, I, P, P, P],
    "vls_batch_autotune_f32": [P, P, I, P, P, P],
    "vls_sdf3_autotune": [P, P, I, P, P],
    "vls_sdf3_autotune_f32": [P, P, I, P, P],
    "vls_sdf3_adjoint_advect_rk3": [P, P, P, P, P, D, P, P, P, P, P, P],
    "vls_sdf3_tangent_advect_rk3": [P, P, P, P, P, P, P, P, P, D, P, P, P],
    "vls_sdf3_adjoint_reinit": [P, P, I, P, P, P],
    "vls_sdf3_tangent_reinit": [P, P, P, I, P, P, P],
}
for _sfx in ("", "_f32"):
    SIGNATURES.update({
        "vls_batch_set_distance" + _sfx: [P, P, P],
        "vls_batch_weno5_rk3_reinit" + _sfx: [P, P, I, P],
        "vls_batch_weno5_rk3_extend" + _sfx: [P, P, P, I, P],
        "vls_batch_advect_rk3" + _sfx: [P, P, P, P, P, P, I, P],
        "vls_batch_normal_velocity" + _sfx: [P] * 7,
        "vls_batch_area_volume_cosine" + _sfx: [P, P, D, P, P],
        "vls_batch_cfl_dt" + _sfx: [P, P, P, P, D, P, P],
        "vls_batch_curl_velocity" + _sfx: [P, P, I, P, P, P, P, P],
        "vls_batch_ellipsoid" + _sfx: [P, P, P, P, P],
        "vls_batch_step" + _sfx: [P, P, P, P, P, P, P, I, I, D, P, P],
    })

SIGNATURES.update({
    "vls_batch_wenork3_reinit": [P, P, I, P],
    "vls_batch_wenork3_extend": [P, P, P, I, P],
    "vls_batch_enork2_extend": [P, P, P, I, P],
})

_lib = None


class VLSError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VLSError(f"HIP library not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        l = ctypes.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(l, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_char_p if name in ("vls_last_error", "vls_version") else ctypes.c_int
        _lib = l
    return _lib


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().vls_last_error().decode()
        raise VLSError(f"{name} failed ({rc}): {msg}")
    return rc


def exported_names():
    return list(SIGNATURES)
