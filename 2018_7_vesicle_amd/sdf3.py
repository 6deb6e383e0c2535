"""Host-side mirror of the reference's level-set objects over the C ABI.

`GD3` and `SDF3` keep the names, argument meaning and results of the
reference's MATLAB classes (src/7_0_paper/2_0_Vesicle/+SD/@GD3/GD3.m,
+SD/@SDF3/*.m) for the hot-path methods.  Fields are torch float64 tensors on
the GPU, contiguous, with shape (pges, cols, rows) = (Nz, Nx, Ny): in memory
that is the reference's column-major [rows, cols, pges] array
(ind = p*rows*cols + c*rows + r).  torch supplies device memory and the
stream; all arithmetic runs in the HIP library.

An SDF3 built on a float32 F runs the FP32-bulk variants of the methods
(the *_f32 entry points): fields are float32, the subcell distances, the
interface points and the reductions are FP64.

The module-level `feval_*` functions are the kernel-level drop-ins: the
reference's feval(kernel, args...) with the same positional arguments, writing
the non-const arguments in place.
"""
import ctypes

import torch

from ._lib import D, P, PD, VLSError, call

__all__ = ["GD3", "SDF3", "VLSError"] + [
    "feval_boundary_location", "feval_boundary_correction", "feval_re_step", "feval_eno2_re_step",
    "feval_upwind_normal", "feval_boundary_interpolant", "feval_extend_step", "feval_advection_step",
    "feval_weno_derivative", "feval_set_calculus_toolbox", "TOOLBOX_FIELDS",
    "feval_auxi_set_calculus_toolbox", "feval_aux_calculus_toolbox", "AUX_TOOLBOX_FIELDS", "AUX_METHOD_FIELDS",
    "TENSION_PRESSURE_FIELDS",
]

TOOLBOX_FIELDS = ["Fx", "Fy", "Fz", "FGradMag", "Nx", "Ny", "Nz", "Fxx", "Fyy", "Fzz", "Fxy", "Fyz", "Fzx",
                  "FLaplacian", "MeanCurvature", "GaussianCurvature", "Heaviside", "DiracDelta"]
# auxi_set_calculus_toolbox's 27 outputs under the @SDF3 property names
# (+SD/@SDF3/GPUAsetCalculusToolBox.m:40-43)
AUX_TOOLBOX_FIELDS = ["Ax", "Ay", "Az", "AGradMag", "Cx", "Cy", "Cz", "NormCrossAF", "tx", "ty", "tz", "nx", "ny",
                      "nz", "Axx", "Ayy", "Azz", "Axy", "Ayz", "Azx", "ALaplacian", "GeodesicCurvature",
                      "NormalCurvature", "GeodesicTorsion", "BPerpendicular", "AHeaviside", "ADiracDelta"]
# the device scalars of SDF3.tensionPressure, in order
TENSION_PRESSURE_FIELDS = ["Tension", "Pressure", "Dt", "c11", "c12", "s1", "s2", "area", "volume", "vmax"]
# what GPUAsetCalculusToolBox.m sets, in vls_aux_calculus_toolbox's out30 order
AUX_METHOD_FIELDS = ["AHPrimal"] + AUX_TOOLBOX_FIELDS + ["AFDiracDelta", "ADiracDeltaDn"]


def _stream():
    return P(torch.cuda.current_stream().cuda_stream)


def _ptr(t, dtype=torch.float64):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise VLSError("expected a CUDA tensor")
    if t.dtype != dtype:
        raise VLSError(f"expected dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise VLSError("expected a contiguous tensor")
    return P(t.data_ptr())


def _dims(t):
    if t.dim() != 3:
        raise VLSError("fields are 3-D tensors of shape (pges, cols, rows)")
    pges, cols, rows = t.shape
    return rows, cols, pges


class GD3:
    """Uniform grid (+SD/@GD3/GD3.m:86-105).  Built from the three coordinate
    vectors of MATLAB's meshgrid(xv, yv, zv)."""

    def __init__(self, xv, yv, zv, device="cuda"):
        self.xv = torch.as_tensor(xv, dtype=torch.float64)
        self.yv = torch.as_tensor(yv, dtype=torch.float64)
        self.zv = torch.as_tensor(zv, dtype=torch.float64)
        self.device = torch.device(device)
        self.Dx = float(self.xv[1] - self.xv[0])
        self.Dy = float(self.yv[1] - self.yv[0])
        self.Dz = float(self.zv[1] - self.zv[0])
        self.Ds = (self.Dx + self.Dy + self.Dz) / 3
        self.EltVol = self.Dx * self.Dy * self.Dz
        self.mrows, self.ncols, self.lshts = len(self.yv), len(self.xv), len(self.zv)
        self.NumElt = self.mrows * self.ncols * self.lshts
        self.Size = (self.lshts, self.ncols, self.mrows)

    def meshgrid(self):
        """X, Y, Z of MATLAB meshgrid as (pges, cols, rows) tensors on the device."""
        Z, X, Y = torch.meshgrid(self.zv, self.xv, self.yv, indexing="ij")
        return (X.contiguous().to(self.device), Y.contiguous().to(self.device), Z.contiguous().to(self.device))

    def zeros(self, dtype=torch.float64):
        return torch.zeros(self.Size, dtype=dtype, device=self.device)

    # the smoothers of the vesicle step (GD3.m:435-486), FP64, periodic box
    def _field(self, t):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or tuple(t.shape) != self.Size:
            raise VLSError(f"expected a float64 field of shape {self.Size}")
        return _ptr(t)

    def _smooth(self, name, rhs, Dt, Alpha, diffusion, out=None):
        out = torch.empty_like(rhs) if out is None else out
        call(name, self._field(out), self._field(rhs), self.mrows, self.ncols, self.lshts, self.Dx, self.Dy, self.Dz,
             float(Dt), float(Alpha), int(diffusion), _stream())
        return out

    def smoothGMRES(self, rhs, Dt, Alpha, out=None):
        """(Idt + Alpha*Dt*LBiLaplacian)^-1 rhs (GD3.m:435-451): the solution of the
        system MATLAB's gmres solves to 1e-12, by the DFT that diagonalizes it."""
        return self._smooth("vls_smooth_gmres", rhs, Dt, Alpha, 0, out)

    def smoothFFT(self, rhs, Dt, Alpha, out=None):
        """GD3.m:453-458: divide by the spectral symbol 1 + Alpha*Dt*|k|^4."""
        return self._smooth("vls_smooth_fft", rhs, Dt, Alpha, 0, out)

    def smoothDiffusionGMRES(self, rhs, Dt, Alpha, out=None):
        """(Idt - Alpha*Dt*LLaplacian)^-1 rhs (GD3.m:461-477), as smoothGMRES."""
        return self._smooth("vls_smooth_gmres", rhs, Dt, Alpha, 1, out)

    def smoothDiffusionFFT(self, rhs, Dt, Alpha, out=None):
        """GD3.m:479-484: divide by 1 + Alpha*Dt*|k|^2."""
        return self._smooth("vls_smooth_fft", rhs, Dt, Alpha, 1, out)

    def smoothOperator(self, x, Dt, Alpha, diffusion=False, out=None):
        """Op*x of the smoothGMRES systems: (Idt + Alpha*Dt*LBiLaplacian) x, or
        (Idt - Alpha*Dt*LLaplacian) x with diffusion (GD3.m:438, 464)."""
        out = torch.empty_like(x) if out is None else out
        call("vls_smooth_operator", self._field(out), self._field(x), self.mrows, self.ncols, self.lshts, self.Dx,
             self.Dy, self.Dz, float(Dt), float(Alpha), int(bool(diffusion)), _stream())
        return out

    def bendingSpeed(self, H, G, Kappa=1.0, SponC=0.0, out=None):
        """NormalSpeedBend of p00_test_singlephase_relax.m:79-83 from the mean
        curvature H and Gaussian curvature G, periodic Laplacian."""
        out = torch.empty_like(H) if out is None else out
        call("vls_bending_speed", self._field(out), self._field(H), self._field(G), self.mrows, self.ncols,
             self.lshts, self.Dx, self.Dy, self.Dz, float(Kappa), float(SponC), _stream())
        return out


class SDF3:
    """Signed-distance level set (+SD/@SDF3/SDF3.m) with its hot-path methods.

    Owns the device state of the MATLAB object (xpr..zpd, min_dist and the
    work fields of the iteration loops) through one C-ABI handle."""

    def __init__(self, grid, F):
        if F.dtype not in (torch.float64, torch.float32):
            raise VLSError(f"F must be float64 or float32, got {F.dtype}")
        self.dtype = F.dtype
        self.GD3 = grid
        self._h = ctypes.c_void_p()
        call("vls_sdf3_create", ctypes.byref(self._h), grid.mrows, grid.ncols, grid.lshts, grid.Dx, grid.Dy,
             grid.Dz)
        self.F = F
        for name in TOOLBOX_FIELDS:
            setattr(self, name, None)
        self.HPrimal = None
        # the auxiliary level set A (the @SDF3 property the AsetCalculusToolBox
        # methods read) and its toolbox
        self.A = None
        for name in AUX_METHOD_FIELDS:
            setattr(self, name, None)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize()
                call("vls_sdf3_destroy", h)
            except Exception:
                pass
            self._h = None

    def _check(self, t):
        if tuple(t.shape) != self.GD3.Size:
            raise VLSError(f"field shape {tuple(t.shape)} != grid {self.GD3.Size}")
        return _ptr(t, self.dtype)

    def _fn(self, name):
        """the entry point of `name` for this object's field type"""
        return name + "_f32" if self.dtype == torch.float32 else name

    # -- launch shape ------------------------------------------------------------
    def setTileRows(self, tile_rows):
        """tile shape of the bulk stage kernels: 64 (64 x 4, default) or 32 (32 x 8)"""
        call("vls_sdf3_set_tile_rows", self._h, int(tile_rows))

    def setZmarchDepth(self, depth):
        """pages per block of the z-marching bulk stage kernels (0: chosen by the box size)"""
        call("vls_sdf3_set_zmarch_depth", self._h, int(depth))

    def setReinitFP64Tail(self, iterations):
        """FP32 objects: the last `iterations` iterations of each reinitialize_
        call run in FP64 with the call's mask and distances (0: none)"""
        call("vls_sdf3_set_reinit_fp64_tail", self._h, int(iterations))

    def launch(self):
        """(tile rows, z-march depth) the bulk stage kernels launch with"""
        tr, zd = ctypes.c_int(), ctypes.c_int()
        call("vls_sdf3_get_launch", self._h, ctypes.byref(tr), ctypes.byref(zd))
        return tr.value, zd.value

    def autotune(self, iterations=2):
        """time extension and reinitialization stages off self.F's interface with
        each tile shape and z-march depth and keep the fastest (the sum of both);
        leaves the distances of self.F set.  Results do not depend on the
        choice.  Returns the tile rows."""
        tr = ctypes.c_int()
        call(self._fn("vls_sdf3_autotune"), self._h, self._check(self.F), int(iterations), ctypes.byref(tr),
             _stream())
        return tr.value

    # -- setDistance (+SD/@SDF3/setDistance.m:3-21) --------------------------
    def setDistance(self):
        call(self._fn("vls_sdf3_set_distance"), self._h, self._check(self.F), _stream())

    def distances(self):
        """copies of (xpr, xpl, ypf, ypb, zpu, zpd, min_dist) of the last setDistance"""
        ptrs = (ctypes.c_void_p * 7)()
        call("vls_sdf3_distances", self._h, ptrs)
        out = []
        n = self.GD3.NumElt
        for k in range(7):
            t = self.GD3.zeros()
            _hip_copy_d2d(t.data_ptr(), ptrs[k], n * 8)
            out.append(t)
        return out

    # -- reinitialization ------------------------------------------------------
    def WENO5RK3Reinitialization(self, F, iteration):
        """+SD/@SDF3/WENO5RK3Reinitialization.m:3-43: mask = F<0, distances of the last setDistance."""
        return self.reinitialize_(F.clone(), iteration)

    # In-place forms of the methods: the argument is updated and returned.  A
    # call is captured into a graph keyed by its pointers, so a loop that
    # passes the same tensors replays one graph.
    def reinitialize_(self, F, iteration):
        call(self._fn("vls_sdf3_weno5_rk3_reinit"), self._h, self._check(F), int(iteration), _stream())
        return F

    def extend_(self, C, iteration):
        call(self._fn("vls_sdf3_weno5_rk3_extend"), self._h, self._check(self.F), self._check(C), int(iteration),
             _stream())
        return C

    def redistance_(self, A, iter1, iter2):
        call(self._fn("vls_sdf3_surface_redistance"), self._h, self._check(self.F), self._check(A), int(iter1),
             int(iter2), _stream())
        return A

    def advect_(self, F, vx, vy, vz, dt, steps=1):
        call(self._fn("vls_sdf3_advect_rk3"), self._h, self._check(F), self._check(vx), self._check(vy),
             self._check(vz), float(dt), int(steps), _stream())
        return F

    def ENORK2Reinitialization(self, F, iteration, rk2=True, cfl=0.3):
        """+SD/@SDF3/ENORK2Reinitialization.m:3-45 (rk2=True, cfl=0.3 is the reference;
        rk2=False is forward Euler).  FP64 only."""
        if self.dtype != torch.float64:
            raise VLSError("ENORK2Reinitialization is FP64 only")
        NewF = F.clone()
        call("vls_sdf3_eno2_reinit", self._h, self._check(NewF), int(iteration), 1 if rk2 else 0, float(cfl),
             _stream())
        return NewF

    def WENORK3Reinitialization(self, F, iteration):
        """+SD/@SDF3/WENORK3Reinitialization.m:3-61: cubic distances of F, mask = F<0,
        RK3 of the WENO-RK3 re_step.  FP64 only."""
        if self.dtype != torch.float64:
            raise VLSError("WENORK3Reinitialization is FP64 only")
        NewF = F.clone()
        call("vls_sdf3_wenork3_reinit", self._h, self._check(NewF), int(iteration), _stream())
        return NewF

    # -- extension and surface redistance -------------------------------------
    def WENO5RK3Extend(self, C, iteration):
        """+SD/@SDF3/WENO5RK3Extend.m:9-76 (self.F and its last setDistance define the interface)."""
        return self.extend_(C.clone(), iteration)

    def ENORK2Extend(self, C, iteration):
        """+SD/@SDF3/ENORK2Extend.m:9-81: ENO2 extension of C off the zero set of self.F
        (its own quadratic-ENO distances), Heun RK2.  FP64 only."""
        return self._variant_extend("vls_sdf3_enork2_extend", "ENORK2Extend", C, iteration)

    def WENORK3Extend(self, C, iteration):
        """+SD/@SDF3/WENORK3Extend.m:9-98: WENO-RK3 extension of C off the zero set of
        self.F (cubic distances and interface values), Shu-Osher RK3.  FP64 only."""
        return self._variant_extend("vls_sdf3_wenork3_extend", "WENORK3Extend", C, iteration)

    def _variant_extend(self, fn, what, C, iteration):
        if self.dtype != torch.float64:
            raise VLSError(f"{what} is FP64 only")
        NewC = C.clone()
        call(fn, self._h, self._check(self.F), self._check(NewC), int(iteration), _stream())
        return NewC

    def WENO5RK3ClosetPointSurfaceRedistance(self, A, iter1, iter2):
        """+SD/@SDF3/WENO5RK3ClosetPointSurfaceRedistance.m:4-157"""
        return self.redistance_(A.clone(), iter1, iter2)

    # -- advection -------------------------------------------------------------
    def WENO5RK3Advection(self, F, vx, vy, vz, dt, steps=1):
        """TVD-RK3 (Shu-Osher form of +SD/@SDF3/WENORK3SurfaceConservationLaw.m:10-29) of
        F_t + v.grad F = 0 with the WENO5 operator of advection_step.cu:30-56."""
        return self.advect_(F.clone(), vx, vy, vz, dt, steps)

    def maxSpeed(self, vx, vy, vz):
        out = ctypes.c_double()
        call(self._fn("vls_sdf3_max_speed"), self._h, self._check(vx), self._check(vy), self._check(vz),
             ctypes.byref(out), _stream())
        return out.value

    # -- calculus toolbox, area and volume --------------------------------------
    def GPUsetCalculusToolBox(self):
        """+SD/@SDF3/GPUsetCalculusToolBox.m:2-40 (HPrimal = max(F,0)).  FP64 only."""
        if self.dtype != torch.float64:
            raise VLSError("GPUsetCalculusToolBox is FP64 only")
        g = self.GD3
        self.HPrimal = torch.clamp(self.F, min=0.0).contiguous()
        outs = [g.zeros() for _ in TOOLBOX_FIELDS]
        feval_set_calculus_toolbox(outs, self.F, self.HPrimal, g.mrows, g.ncols, g.lshts, g.Dx, g.Dy, g.Dz, g.Ds)
        for name, t in zip(TOOLBOX_FIELDS, outs):
            setattr(self, name, t)

    def GPUAsetCalculusToolBox(self):
        """+SD/@SDF3/GPUAsetCalculusToolBox.m:2-64: the geometry of the auxiliary
        field A on the surface of F -- AHPrimal = max(A,0), the 27 fields of
        auxi_set_calculus_toolbox, AFDiracDelta and ADiracDeltaDn (the last two
        with @GD3's periodic differences).  Needs self.A and the toolbox of F
        (GPUsetCalculusToolBox) first.  FP64 only."""
        if self.dtype != torch.float64:
            raise VLSError("GPUAsetCalculusToolBox is FP64 only")
        if self.A is None:
            raise VLSError("GPUAsetCalculusToolBox needs the auxiliary field A")
        if self.Fx is None:
            raise VLSError("GPUAsetCalculusToolBox needs GPUsetCalculusToolBox first")
        g = self.GD3
        outs = [g.zeros() for _ in AUX_METHOD_FIELDS]
        fin = [getattr(self, n) for n in TOOLBOX_FIELDS[:13]] + [self.Heaviside]
        for t in fin + [self.A]:
            self._check(t)
        feval_aux_calculus_toolbox(outs, self.A, fin, g.mrows, g.ncols, g.lshts, g.Dx, g.Dy, g.Dz, g.Ds)
        for name, t in zip(AUX_METHOD_FIELDS, outs):
            setattr(self, name, t)

    def tensionPressure(self, H, Vb, InitialArea, InitialVolume, Dt=None, CFLNumber=None, out=None):
        """The tension/pressure step of p00_test_singlephase_relax.m:67-109 on the
        toolbox of F (GPUsetCalculusToolBox first): H the extended mean
        curvature, Vb the extended bending speed.  Dt = CFLNumber*Ds/max|Vb|
        on |F|<2Ds unless Dt is given.  Returns (normalSpeed, info) with info a
        device tensor {Tension, Pressure, Dt, c11, c12, s1, s2, area, volume,
        max|Vb|}; nothing is read back to the host."""
        if self.dtype != torch.float64:
            raise VLSError("tensionPressure is FP64 only")
        if self.DiracDelta is None:
            raise VLSError("tensionPressure needs GPUsetCalculusToolBox first")
        if (Dt is None or Dt <= 0) and (CFLNumber is None or CFLNumber <= 0):
            raise VLSError("give Dt > 0 or CFLNumber > 0")
        speed = self.GD3.zeros() if out is None else out
        info = torch.zeros(10, dtype=torch.float64, device=self.F.device)
        call("vls_sdf3_tension_pressure", self._h, self._check(self.F), self._check(H), self._check(Vb),
             self._check(self.DiracDelta), self._check(self.FGradMag), self._check(self.Heaviside),
             float(InitialArea), float(InitialVolume), float(Dt or 0.0), float(CFLNumber or 0.0),
             self._check(speed), _ptr(info), _stream())
        return speed, info

    def calAreaVolume(self):
        """(calArea, calVolume) of +SD/@SDF3/SDF3.m:73-80 with the Towers delta/H,
        fused with the toolbox stencil (no toolbox fields are written)."""
        if self.dtype == torch.float32:
            return self._area_volume_f32(0, 0.0)
        a, v = ctypes.c_double(), ctypes.c_double()
        call("vls_sdf3_area_volume_towers", self._h, self._check(self.F), ctypes.byref(a), ctypes.byref(v),
             _stream())
        return a.value, v.value

    def _area_volume_f32(self, kind, eps):
        out = torch.empty(2, dtype=torch.float64, device=self.F.device)
        call("vls_sdf3_area_volume_device_f32", self._h, self._check(self.F), int(kind), float(eps), _ptr(out),
             _stream())
        a, v = out.tolist()
        return a, v

    def areaVolumeDevice(self, out2, kind=1, eps=None):
        """(area, volume) written to the float64 device tensor out2[0:2] without a
        host synchronization; kind 0 = Towers delta/H, 1 = cosine (eps = 1.5*Dx)."""
        if eps is None:
            eps = 1.5 * self.GD3.Dx
        call(self._fn("vls_sdf3_area_volume_device"), self._h, self._check(self.F), int(kind), float(eps),
             _ptr(out2), _stream())

    def normalVelocity(self, vx, vy, vz, out):
        """u.N with N = grad F/max(|grad F|, 1e-14) (central differences), into `out`"""
        call(self._fn("vls_sdf3_normal_velocity"), self._h, self._check(self.F), self._check(vx),
             self._check(vy), self._check(vz), self._check(out), _stream())
        return out

    def calArea(self):
        return self.calAreaVolume()[0]

    def calVolume(self):
        return self.calAreaVolume()[1]

    def surfaceIntegral(self, field):
        """+SD/@SDF3/SDF3.m:82-85 (needs GPUsetCalculusToolBox first)."""
        mask = self.F.abs() < 2 * self.GD3.Ds
        return float((field[mask] * self.DiracDelta[mask] * self.FGradMag[mask]).sum()) * self.GD3.Ds ** 3

    def cosineAreaVolume(self, eps=None):
        """SurfaceIntegral(1), VolumeIntegral(1) with the cosine-smoothed delta/H of
        src/0_0_setup/+SD/@SDF3/SDF3.m:126-136,178-187; eps defaults to 1.5*Dx (:59)."""
        if eps is None:
            eps = 1.5 * self.GD3.Dx
        if self.dtype == torch.float32:
            return self._area_volume_f32(1, eps)
        a, v = ctypes.c_double(), ctypes.c_double()
        call("vls_sdf3_area_volume_cosine", self._h, self._check(self.F), float(eps), ctypes.byref(a),
             ctypes.byref(v), _stream())
        return a.value, v.value


_hip = None


def _hip_copy_d2d(dst, src, nbytes):
    """device-to-device copy on torch's current stream through the process's HIP runtime"""
    global _hip
    if _hip is None:
        import os

        for line in open("/proc/self/maps"):
            path = line.split()[-1]
            if "libamdhip64" in os.path.basename(path):
                _hip = ctypes.CDLL(path)
                break
        else:
            _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpyAsync.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int, P]
    err = _hip.hipMemcpyAsync(P(dst), P(src), nbytes, 3, _stream())  # 3 = hipMemcpyDeviceToDevice
    if err != 0:
        raise VLSError(f"hipMemcpyAsync failed with {err}")


# ---------------------------------------------------------------------------
# kernel-level drop-ins (feval mirrors)
# ---------------------------------------------------------------------------
def feval_boundary_location(xpr, xpl, ypf, ypb, zpu, zpd, lsf, num_ele, rows, cols, pges, dx, dy, dz):
    call("vls_boundary_location", *[_ptr(t) for t in (xpr, xpl, ypf, ypb, zpu, zpd, lsf)], num_ele, rows, cols,
         pges, dx, dy, dz, _stream())
    return xpr, xpl, ypf, ypb, zpu, zpd


def feval_boundary_correction(xpr, xpl, ypf, ypb, zpu, zpd, lsf, num_ele, rows, cols, pges, dx, dy, dz):
    call("vls_boundary_correction", *[_ptr(t) for t in (xpr, xpl, ypf, ypb, zpu, zpd, lsf)], num_ele, rows, cols,
         pges, dx, dy, dz, _stream())
    return xpr, xpl, ypf, ypb, zpu, zpd


def feval_re_step(step, lsf, mask, deltat, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, num_ele):
    call("vls_re_step", _ptr(step), _ptr(lsf), _ptr(mask, torch.uint8), _ptr(deltat),
         *[_ptr(t) for t in (xpr, xpl, ypf, ypb, zpu, zpd)], rows, cols, pges, dx, dy, dz, num_ele, _stream())
    return step


def feval_eno2_re_step(step, lsf, mask, deltat, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz,
                       num_ele):
    call("vls_eno2_re_step", _ptr(step), _ptr(lsf), _ptr(mask, torch.uint8), _ptr(deltat),
         *[_ptr(t) for t in (xpr, xpl, ypf, ypb, zpu, zpd)], rows, cols, pges, dx, dy, dz, num_ele, _stream())
    return step


def feval_upwind_normal(nx, ny, nz, lsf, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, num_ele):
    call("vls_upwind_normal", *[_ptr(t) for t in (nx, ny, nz, lsf, xpr, xpl, ypf, ypb, zpu, zpd)], rows, cols,
         pges, dx, dy, dz, num_ele, _stream())
    return nx, ny, nz


def feval_boundary_interpolant(cpr, cpl, cpf, cpb, cpu, cpd, xpr, xpl, ypf, ypb, zpu, zpd, lsf, rows, cols, pges,
                               dx, dy, dz, num_ele):
    call("vls_boundary_interpolant",
         *[_ptr(t) for t in (cpr, cpl, cpf, cpb, cpu, cpd, xpr, xpl, ypf, ypb, zpu, zpd, lsf)], rows, cols, pges,
         dx, dy, dz, num_ele, _stream())
    return cpr, cpl, cpf, cpb, cpu, cpd


def feval_extend_step(step, deltat, lsf, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cpr, cpl, cpf, cpb, cpu, cpd,
                      rows, cols, pges, dx, dy, dz, num_ele):
    call("vls_extend_step", *[_ptr(t) for t in (step, deltat, lsf, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cpr,
                                                 cpl, cpf, cpb, cpu, cpd)], rows, cols, pges, dx, dy, dz, num_ele,
         _stream())
    return step


# WENO-RK3 / ENO-RK2 variants (cubic_boundary_correction, WENORK3_*, ENORK2_* of
# +SD/@SDF3/GPUInitialize.m:32-42, 64-102)
def feval_cubic_boundary_correction(xpr, xpl, ypf, ypb, zpu, zpd, lsf, num_ele, rows, cols, pges, dx, dy, dz):
    call("vls_cubic_boundary_correction", *[_ptr(t) for t in (xpr, xpl, ypf, ypb, zpu, zpd, lsf)], num_ele, rows,
         cols, pges, dx, dy, dz, _stream())
    return xpr, xpl, ypf, ypb, zpu, zpd


def feval_wenork3_re_step(step, lsf, mask, deltat, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz,
                          num_ele):
    call("vls_wenork3_re_step", _ptr(step), _ptr(lsf), _ptr(mask, torch.uint8), _ptr(deltat),
         *[_ptr(t) for t in (xpr, xpl, ypf, ypb, zpu, zpd)], rows, cols, pges, dx, dy, dz, num_ele, _stream())
    return step


def _variant_normal(name, nx, ny, nz, lsf, xpr, xpl, ypf, ypb, zpu, zpd, rows, cols, pges, dx, dy, dz, num_ele):
    call(name, *[_ptr(t) for t in (nx, ny, nz, lsf, xpr, xpl, ypf, ypb, zpu, zpd)], rows, cols, pges, dx, dy, dz,
         num_ele, _stream())
    return nx, ny, nz


def feval_enork2_upwind_normal(*args):
    return _variant_normal("vls_enork2_upwind_normal", *args)


def feval_wenork3_upwind_normal(*args):
    return _variant_normal("vls_wenork3_upwind_normal", *args)


def _variant_interpolate(name, cpr, cpl, cpf, cpb, cpu, cpd, xpr, xpl, ypf, ypb, zpu, zpd, lsf, rows, cols, pges,
                         dx, dy, dz, num_ele):
    call(name, *[_ptr(t) for t in (cpr, cpl, cpf, cpb, cpu, cpd, xpr, xpl, ypf, ypb, zpu, zpd, lsf)], rows, cols,
         pges, dx, dy, dz, num_ele, _stream())
    return cpr, cpl, cpf, cpb, cpu, cpd


def feval_enork2_boundary_interpolate(*args):
    return _variant_interpolate("vls_enork2_boundary_interpolate", *args)


def feval_wenork3_boundary_interpolate(*args):
    return _variant_interpolate("vls_wenork3_boundary_interpolate", *args)


def _variant_extend_step(name, step, deltat, lsf, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cpr, cpl, cpf, cpb,
                         cpu, cpd, rows, cols, pges, dx, dy, dz, num_ele):
    call(name, *[_ptr(t) for t in (step, deltat, lsf, vx, vy, vz, xpr, xpl, ypf, ypb, zpu, zpd, cpr, cpl, cpf, cpb,
                                   cpu, cpd)], rows, cols, pges, dx, dy, dz, num_ele, _stream())
    return step


def feval_enork2_extend_step(*args):
    return _variant_extend_step("vls_enork2_extend_step", *args)


def feval_wenork3_extend_step(*args):
    return _variant_extend_step("vls_wenork3_extend_step", *args)


def feval_advection_step(astep, vx, vy, vz, lsf, rows, cols, pges, dx, dy, dz):
    call("vls_advection_step", *[_ptr(t) for t in (astep, vx, vy, vz, lsf)], rows, cols, pges, dx, dy, dz,
         _stream())
    return astep


def feval_weno_derivative(back_x, fore_x, back_y, fore_y, back_z, fore_z, lsf, rows, cols, pges, dx, dy, dz):
    call("vls_weno_derivative", *[_ptr(t) for t in (back_x, fore_x, back_y, fore_y, back_z, fore_z, lsf)], rows,
         cols, pges, dx, dy, dz, _stream())
    return back_x, fore_x, back_y, fore_y, back_z, fore_z


def feval_auxi_set_calculus_toolbox(out27, lsf, ahprimal, fin13, rows, cols, pges, dx, dy, dz, ds):
    """feval(auxi_set_calculus_toolbox, ...) of +SD/@SDF3/GPUAsetCalculusToolBox.m:44-54:
    out27 in AUX_TOOLBOX_FIELDS order; fin13 = Fx..Fzx of F (TOOLBOX_FIELDS[:13])."""
    if len(out27) != 27 or len(fin13) != 13:
        raise VLSError("auxi_set_calculus_toolbox writes 27 fields from 13 fields of F")
    arr = (ctypes.c_void_p * 27)(*[_ptr(t).value for t in out27])
    farr = (ctypes.c_void_p * 13)(*[_ptr(t).value for t in fin13])
    call("vls_auxi_set_calculus_toolbox", arr, _ptr(lsf), _ptr(ahprimal), farr, rows, cols, pges, dx, dy, dz, ds,
         rows * cols * pges, _stream())
    return out27


def feval_aux_calculus_toolbox(out30, A, ftool14, rows, cols, pges, dx, dy, dz, ds):
    """GPUAsetCalculusToolBox.m:2-64 on caller buffers: out30 in AUX_METHOD_FIELDS
    order; ftool14 = Fx..Fzx of F (TOOLBOX_FIELDS[:13]) and F's Heaviside."""
    if len(out30) != 30 or len(ftool14) != 14:
        raise VLSError("the auxiliary toolbox writes 30 fields from 14 fields of F")
    arr = (ctypes.c_void_p * 30)(*[_ptr(t).value for t in out30])
    farr = (ctypes.c_void_p * 14)(*[_ptr(t).value for t in ftool14])
    call("vls_aux_calculus_toolbox", arr, _ptr(A), farr, rows, cols, pges, dx, dy, dz, ds, _stream())
    return out30


def feval_set_calculus_toolbox(out18, lsf, hprimal, rows, cols, pges, dx, dy, dz, ds):
    if len(out18) != 18:
        raise VLSError("set_calculus_toolbox writes 18 fields")
    arr = (ctypes.c_void_p * 18)(*[_ptr(t).value for t in out18])
    call("vls_set_calculus_toolbox", arr, _ptr(lsf), _ptr(hprimal), rows, cols, pges, dx, dy, dz, ds,
         rows * cols * pges, _stream())
    return out18
