"""Launcher for the reference's own kernels (TEST INFRASTRUCTURE ONLY).

`make -C oracle ref` compiles the reference's CUDA sources, unmodified, into
one gfx950 device code object per .cu file under oracle/_ref/.  This module
loads those objects by entry name and launches them with the reference's
launch geometry, which is what +SD/@SDF3/GPUInitialize.m does with
parallel.gpu.CUDAKernel objects:

    ThreadBlockSize = [mrows, floor(256/mrows), 1]          (GPUInitialize.m:7)
    GridSize = [ceil(mrows/TB1), ceil(ncols/TB2), lshts]    (GPUInitialize.m:8-10)

Arguments follow feval semantics: non-const pointers are in/out and are
written in place.  It is used only by tests/ and tests/golden/make_golden.py
to check the CPU oracle and the product; the product never imports it.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")


def available():
    return os.path.exists(os.path.join(REF_DIR, "kernels.txt"))


def _hip_runtime():
    """The HIP runtime torch already loaded (one runtime per process)."""
    import torch  # noqa: F401  (loads torch's libamdhip64)

    torch.cuda.init()
    for line in open("/proc/self/maps"):
        path = line.split()[-1]
        if "libamdhip64" in os.path.basename(path):
            return ctypes.CDLL(path)
    return ctypes.CDLL("libamdhip64.so")


class RefKernels:
    def __init__(self, ref_dir=REF_DIR):
        self.hip = _hip_runtime()
        self.hip.hipModuleLoad.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p]
        self.hip.hipModuleGetFunction.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_char_p]
        self.hip.hipModuleLaunchKernel.argtypes = [ctypes.c_void_p] + [ctypes.c_uint] * 7 + [
            ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]
        self.dir = ref_dir
        self.symbols = {}
        for line in open(os.path.join(ref_dir, "kernels.txt")):
            co, sym = line.split()
            self.symbols[(co, self._demangled_name(sym))] = sym
        self.modules = {}
        self.functions = {}

    @staticmethod
    def _demangled_name(sym):
        # _Z<len><name>... -> name
        assert sym.startswith("_Z")
        i = 2
        j = i
        while sym[j].isdigit():
            j += 1
        n = int(sym[i:j])
        return sym[j:j + n]

    def _function(self, co, entry):
        key = (co, entry)
        if key in self.functions:
            return self.functions[key]
        if co not in self.modules:
            mod = ctypes.c_void_p()
            err = self.hip.hipModuleLoad(ctypes.byref(mod), os.path.join(self.dir, co).encode())
            if err != 0:
                raise RuntimeError(f"hipModuleLoad({co}) failed with {err}")
            self.modules[co] = mod
        fn = ctypes.c_void_p()
        err = self.hip.hipModuleGetFunction(ctypes.byref(fn), self.modules[co], self.symbols[key].encode())
        if err != 0:
            raise RuntimeError(f"hipModuleGetFunction({entry}) failed with {err}")
        self.functions[key] = fn
        return fn

    def launch(self, co, entry, args, rows, cols, pges, stream=None):
        """args: list of ('p', tensor) | ('i', int) | ('d', float)."""
        import torch

        if rows > 256:
            raise ValueError("the reference launch geometry needs rows <= 256 (GPUInitialize.m:7)")
        tb = (rows, 256 // rows, 1)
        grid = ((rows + tb[0] - 1) // tb[0], (cols + tb[1] - 1) // tb[1], pges)
        holders = []
        for kind, val in args:
            if kind == "p":
                assert val.is_cuda and val.is_contiguous()
                holders.append(ctypes.c_void_p(val.data_ptr()))
            elif kind == "i":
                holders.append(ctypes.c_int(int(val)))
            elif kind == "d":
                holders.append(ctypes.c_double(float(val)))
            else:
                raise ValueError(kind)
        params = (ctypes.c_void_p * len(holders))(*[ctypes.cast(ctypes.pointer(h), ctypes.c_void_p) for h in holders])
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        err = self.hip.hipModuleLaunchKernel(self._function(co, entry), grid[0], grid[1], grid[2],
                                             tb[0], tb[1], tb[2], 0, ctypes.c_void_p(stream), params, None)
        if err != 0:
            raise RuntimeError(f"hipModuleLaunchKernel({entry}) failed with {err}")

    # ---- feval mirrors ------------------------------------------------------
    def boundary_location(self, d6, lsf, rows, cols, pges, dx, dy, dz):
        args = [("p", t) for t in d6] + [("p", lsf), ("i", rows * cols * pges), ("i", rows), ("i", cols),
                                         ("i", pges), ("d", dx), ("d", dy), ("d", dz)]
        self.launch("boundary_location.co", "boundary_location", args, rows, cols, pges)

    def boundary_correction(self, d6, lsf, rows, cols, pges, dx, dy, dz):
        args = [("p", t) for t in d6] + [("p", lsf), ("i", rows * cols * pges), ("i", rows), ("i", cols),
                                         ("i", pges), ("d", dx), ("d", dy), ("d", dz)]
        self.launch("boundary_correction.co", "boundary_correction", args, rows, cols, pges)

    def _re_step(self, co, step, lsf, mask, deltat, d6, rows, cols, pges, dx, dy, dz):
        args = [("p", step), ("p", lsf), ("p", mask), ("p", deltat)] + [("p", t) for t in d6] + [
            ("i", rows), ("i", cols), ("i", pges), ("d", dx), ("d", dy), ("d", dz), ("i", rows * cols * pges)]
        self.launch(co, "re_step", args, rows, cols, pges)

    def re_step(self, *a):
        self._re_step("re_step.co", *a)

    def eno2_re_step(self, *a):
        self._re_step("boundary_correction.co", *a)

    def upwind_normal(self, n3, lsf, d6, rows, cols, pges, dx, dy, dz):
        args = [("p", t) for t in n3] + [("p", lsf)] + [("p", t) for t in d6] + [
            ("i", rows), ("i", cols), ("i", pges), ("d", dx), ("d", dy), ("d", dz), ("i", rows * cols * pges)]
        self.launch("upwind_normal.co", "upwind_normal", args, rows, cols, pges)

    def boundary_interpolant(self, cp6, d6, lsf, rows, cols, pges, dx, dy, dz):
        args = [("p", t) for t in cp6] + [("p", t) for t in d6] + [("p", lsf), ("i", rows), ("i", cols),
                                                                  ("i", pges), ("d", dx), ("d", dy), ("d", dz),
                                                                  ("i", rows * cols * pges)]
        self.launch("boundary_interpolant.co", "boundary_interpolant", args, rows, cols, pges)

    def extend_step(self, step, deltat, lsf, v3, d6, cp6, rows, cols, pges, dx, dy, dz):
        args = [("p", step), ("p", deltat), ("p", lsf)] + [("p", t) for t in v3] + [("p", t) for t in d6] + [
            ("p", t) for t in cp6] + [("i", rows), ("i", cols), ("i", pges), ("d", dx), ("d", dy), ("d", dz),
                                      ("i", rows * cols * pges)]
        self.launch("extend_step.co", "extend_step", args, rows, cols, pges)

    def advection_step(self, astep, v3, lsf, rows, cols, pges, dx, dy, dz):
        args = [("p", astep)] + [("p", t) for t in v3] + [("p", lsf), ("i", rows), ("i", cols), ("i", pges),
                                                         ("d", dx), ("d", dy), ("d", dz)]
        self.launch("advection_step.co", "advection_step", args, rows, cols, pges)

    def weno_derivative(self, out6, lsf, rows, cols, pges, dx, dy, dz):
        args = [("p", t) for t in out6] + [("p", lsf), ("i", rows), ("i", cols), ("i", pges),
                                           ("d", dx), ("d", dy), ("d", dz)]
        self.launch("nonoscillating_interpolant.co", "weno_derivative", args, rows, cols, pges)

    def set_calculus_toolbox(self, out18, lsf, hprimal, rows, cols, pges, dx, dy, dz, ds):
        args = [("p", t) for t in out18] + [("p", lsf), ("p", hprimal), ("i", rows), ("i", cols), ("i", pges),
                                            ("d", dx), ("d", dy), ("d", dz), ("d", ds), ("i", rows * cols * pges)]
        self.launch("set_calculus_toolbox.co", "set_calculus_toolbox", args, rows, cols, pges)

    def auxi_set_calculus_toolbox(self, out27, lsf, ahprimal, fin13, rows, cols, pges, dx, dy, dz, ds):
        args = [("p", t) for t in out27] + [("p", lsf), ("p", ahprimal)] + [("p", t) for t in fin13] + [
            ("i", rows), ("i", cols), ("i", pges), ("d", dx), ("d", dy), ("d", dz), ("d", ds),
            ("i", rows * cols * pges)]
        self.launch("auxi_set_calculus_toolbox.co", "auxi_set_calculus_toolbox", args, rows, cols, pges)

    # ---- the WENO-RK3 / ENO-RK2 variants (GPUInitialize.m:32-42, 64-102) ------
    def cubic_boundary_correction(self, d6, lsf, rows, cols, pges, dx, dy, dz):
        args = [("p", t) for t in d6] + [("p", lsf), ("i", rows * cols * pges), ("i", rows), ("i", cols),
                                         ("i", pges), ("d", dx), ("d", dy), ("d", dz)]
        self.launch("weno_rk3_reinitialization.co", "boundary_correction", args, rows, cols, pges)

    def wenork3_re_step(self, *a):
        self._re_step("weno_rk3_reinitialization.co", *a)

    def variant_upwind_normal(self, scheme, n3, lsf, d6, rows, cols, pges, dx, dy, dz):
        args = [("p", t) for t in n3] + [("p", lsf)] + [("p", t) for t in d6] + [
            ("i", rows), ("i", cols), ("i", pges), ("d", dx), ("d", dy), ("d", dz), ("i", rows * cols * pges)]
        self.launch(scheme + "_extend.co", "upwind_normal", args, rows, cols, pges)

    def variant_boundary_interpolate(self, scheme, cp6, d6, lsf, rows, cols, pges, dx, dy, dz):
        args = [("p", t) for t in cp6] + [("p", t) for t in d6] + [("p", lsf), ("i", rows), ("i", cols),
                                                                  ("i", pges), ("d", dx), ("d", dy), ("d", dz),
                                                                  ("i", rows * cols * pges)]
        self.launch(scheme + "_extend.co", "boundary_interpolate", args, rows, cols, pges)

    def variant_extend_step(self, scheme, step, deltat, lsf, v3, d6, cp6, rows, cols, pges, dx, dy, dz):
        args = [("p", step), ("p", deltat), ("p", lsf)] + [("p", t) for t in v3] + [("p", t) for t in d6] + [
            ("p", t) for t in cp6] + [("i", rows), ("i", cols), ("i", pges), ("d", dx), ("d", dy), ("d", dz),
                                      ("i", rows * cols * pges)]
        self.launch(scheme + "_extend.co", "extend_step", args, rows, cols, pges)
