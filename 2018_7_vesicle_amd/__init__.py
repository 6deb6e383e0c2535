"""MI355X-native level-set hot path of tiankuizhang/2018_7_Vesicle.

HIP kernels for gfx950 behind the C ABI of include/vesicle_ls.h, with a host
mirror of the reference's @GD3/@SDF3 MATLAB classes (sdf3.py).  The directory
name starts with a digit, so import it with
importlib.import_module("2018_7_vesicle_amd").
"""
from ._lib import LIB_PATH, VLSError, exported_names, lib  # noqa: F401
from .sdf3 import (AUX_METHOD_FIELDS, AUX_TOOLBOX_FIELDS, GD3, TENSION_PRESSURE_FIELDS, SDF3, TOOLBOX_FIELDS, feval_advection_step, feval_boundary_correction,  # noqa: F401
                   feval_boundary_interpolant, feval_boundary_location, feval_cubic_boundary_correction,
                   feval_eno2_re_step, feval_enork2_boundary_interpolate, feval_enork2_extend_step,
                   feval_enork2_upwind_normal, feval_extend_step, feval_re_step, feval_set_calculus_toolbox,
                   feval_upwind_normal, feval_weno_derivative, feval_auxi_set_calculus_toolbox, feval_aux_calculus_toolbox, feval_wenork3_boundary_interpolate,
                   feval_wenork3_extend_step, feval_wenork3_re_step, feval_wenork3_upwind_normal)
from . import adjoint, matio, pipeline, relax  # noqa: F401
from .batch import Batch, BatchConfig, BatchPipeline, set_graphs  # noqa: F401

__version__ = "0.1.0"
