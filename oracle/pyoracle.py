"""numpy front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Wraps oracle/liblsoracle.so (built by `make -C oracle`).  Arrays are numpy
float64 in C order with shape (pges, cols, rows) = (Nz, Nx, Ny), which is the
reference's column-major [rows, cols, pges] layout: ind = p*rows*cols + c*rows + r.
Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsoracle.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def _dims(a):
    pges, cols, rows = a.shape
    return rows, cols, pges


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_num_threads(n):
    lib().or_set_num_threads(_I(int(n)))


def num_threads():
    return int(lib().or_num_threads())


# --- kernels (feval mirrors; in/out arrays are modified in place) ----------

def boundary_location(d6, lsf, dx, dy, dz, order=6):
    rows, cols, pges = _dims(lsf)
    fn = {6: lib().or_boundary_location, 2: lib().or_boundary_correction,
          3: lib().or_cubic_boundary_correction}[order]
    fn(*[_p(a) for a in d6], _p(lsf), _I(lsf.size), _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz))


def prefilled_distances(shape, dx, dy, dz):
    return [np.full(shape, v, dtype=np.float64) for v in (dx, dx, dy, dy, dz, dz)]


def re_step(lsf, mask, deltat, d6, dx, dy, dz, scheme="weno5"):
    rows, cols, pges = _dims(lsf)
    step = np.zeros_like(lsf)
    fn = {"weno5": lib().or_re_step, "eno2": lib().or_eno2_re_step, "wenork3": lib().or_wenork3_re_step}[scheme]
    fn(_p(step), _p(lsf), _p(np.ascontiguousarray(mask, dtype=np.uint8)), _p(deltat), *[_p(a) for a in d6],
       _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz), _I(lsf.size))
    return step


def upwind_normal(lsf, d6, dx, dy, dz):
    rows, cols, pges = _dims(lsf)
    out = [np.zeros_like(lsf) for _ in range(3)]
    lib().or_upwind_normal(*[_p(a) for a in out], _p(lsf), *[_p(a) for a in d6], _I(rows), _I(cols), _I(pges),
                           _D(dx), _D(dy), _D(dz), _I(lsf.size))
    return out


def boundary_interpolant(d6, C, dx, dy, dz):
    rows, cols, pges = _dims(C)
    out = [np.zeros_like(C) for _ in range(6)]
    lib().or_boundary_interpolant(*[_p(a) for a in out], *[_p(a) for a in d6], _p(C), _I(rows), _I(cols),
                                  _I(pges), _D(dx), _D(dy), _D(dz), _I(C.size))
    return out


def extend_step(deltat, C, v3, d6, cp6, dx, dy, dz):
    rows, cols, pges = _dims(C)
    step = np.zeros_like(C)
    lib().or_extend_step(_p(step), _p(deltat), _p(C), *[_p(a) for a in v3], *[_p(a) for a in d6],
                         *[_p(a) for a in cp6], _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz), _I(C.size))
    return step


def variant_upwind_normal(lsf, d6, dx, dy, dz, scheme):
    """upwind_normal of the ENO-RK2 ("enork2") or WENO-RK3 ("wenork3") extension"""
    rows, cols, pges = _dims(lsf)
    out = [np.zeros_like(lsf) for _ in range(3)]
    fn = {"enork2": lib().or_enork2_upwind_normal, "wenork3": lib().or_wenork3_upwind_normal}[scheme]
    fn(*[_p(a) for a in out], _p(lsf), *[_p(a) for a in d6], _I(rows), _I(cols), _I(pges), _D(dx), _D(dy),
       _D(dz), _I(lsf.size))
    return out


def variant_boundary_interpolate(d6, C, dx, dy, dz, scheme):
    rows, cols, pges = _dims(C)
    out = [np.zeros_like(C) for _ in range(6)]
    fn = {"enork2": lib().or_enork2_boundary_interpolate, "wenork3": lib().or_wenork3_boundary_interpolate}[scheme]
    fn(*[_p(a) for a in out], *[_p(a) for a in d6], _p(C), _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz),
       _I(C.size))
    return out


def variant_extend_step(deltat, C, v3, d6, cp6, dx, dy, dz, scheme):
    rows, cols, pges = _dims(C)
    step = np.zeros_like(C)
    fn = {"enork2": lib().or_enork2_extend_step, "wenork3": lib().or_wenork3_extend_step}[scheme]
    fn(_p(step), _p(deltat), _p(C), *[_p(a) for a in v3], *[_p(a) for a in d6], *[_p(a) for a in cp6], _I(rows),
       _I(cols), _I(pges), _D(dx), _D(dy), _D(dz), _I(C.size))
    return step


def advection_step(v3, lsf, dx, dy, dz):
    rows, cols, pges = _dims(lsf)
    out = np.zeros_like(lsf)
    lib().or_advection_step(_p(out), *[_p(a) for a in v3], _p(lsf), _I(rows), _I(cols), _I(pges),
                            _D(dx), _D(dy), _D(dz))
    return out


def weno_derivative(lsf, dx, dy, dz):
    """returns [back_x, fore_x, back_y, fore_y, back_z, fore_z]"""
    rows, cols, pges = _dims(lsf)
    out = [np.zeros_like(lsf) for _ in range(6)]
    lib().or_weno_derivative(*[_p(a) for a in out], _p(lsf), _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz))
    return out


TOOLBOX_FIELDS = ["Fx", "Fy", "Fz", "FGradMag", "Nx", "Ny", "Nz", "Fxx", "Fyy", "Fzz", "Fxy", "Fyz", "Fzx",
                  "FLaplacian", "MeanCurvature", "GaussianCurvature", "Heaviside", "DiracDelta"]


def set_calculus_toolbox(lsf, hprimal, dx, dy, dz, ds):
    rows, cols, pges = _dims(lsf)
    out = [np.zeros_like(lsf) for _ in range(18)]
    arr = (_P * 18)(*[a.ctypes.data for a in out])
    lib().or_set_calculus_toolbox(arr, _p(lsf), _p(hprimal), _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz),
                                  _D(ds), _I(lsf.size))
    return dict(zip(TOOLBOX_FIELDS, out))


AUX_TOOLBOX_FIELDS = ["Ax", "Ay", "Az", "AGradMag", "Cx", "Cy", "Cz", "NormCrossAF", "tx", "ty", "tz", "nx", "ny",
                      "nz", "Axx", "Ayy", "Azz", "Axy", "Ayz", "Azx", "ALaplacian", "GeodesicCurvature",
                      "NormalCurvature", "GeodesicTorsion", "BPerpendicular", "AHeaviside", "ADiracDelta"]


def auxi_set_calculus_toolbox(A, ahprimal, fin13, dx, dy, dz, ds):
    """or_auxi_set_calculus_toolbox (auxi_set_calculus_toolbox.cu:9-199); fin13 =
    Fx..Fzx of F in TOOLBOX_FIELDS order.  Returns a dict of the 27 fields."""
    rows, cols, pges = _dims(A)
    out = [np.zeros_like(A) for _ in range(27)]
    fin = [np.ascontiguousarray(a, dtype=np.float64) for a in fin13]
    arr = (_P * 27)(*[a.ctypes.data for a in out])
    farr = (_P * 13)(*[a.ctypes.data for a in fin])
    lib().or_auxi_set_calculus_toolbox(arr, _p(A), _p(ahprimal), farr, _I(rows), _I(cols), _I(pges), _D(dx),
                                       _D(dy), _D(dz), _D(ds), _I(A.size))
    return dict(zip(AUX_TOOLBOX_FIELDS, out))


def aux_calculus_toolbox(A, ftool, dx, dy, dz, ds):
    """or_aux_calculus_toolbox (+SD/@SDF3/GPUAsetCalculusToolBox.m:2-64); ftool is
    set_calculus_toolbox's dict of F.  Returns AHPrimal, the 27 fields,
    AFDiracDelta and ADiracDeltaDn by name."""
    rows, cols, pges = _dims(A)
    out = [np.zeros_like(A) for _ in range(30)]
    fin = [np.ascontiguousarray(ftool[k], dtype=np.float64) for k in TOOLBOX_FIELDS[:13] + ["Heaviside"]]
    arr = (_P * 30)(*[a.ctypes.data for a in out])
    farr = (_P * 14)(*[a.ctypes.data for a in fin])
    lib().or_aux_calculus_toolbox(arr, _p(A), farr, _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz), _D(ds))
    return dict(zip(["AHPrimal"] + AUX_TOOLBOX_FIELDS + ["AFDiracDelta", "ADiracDeltaDn"], out))


def weno_onesided(v1, v2, v3, v4, v5):
    f = lib().or_weno_onesided
    f.restype = _D
    return f(_D(v1), _D(v2), _D(v3), _D(v4), _D(v5))


def weno_pair(p7, ds):
    """(d_fore, d_back) of the uniform WENO5 pair (or_weno_pair, Jiang-Peng form)"""
    p = np.ascontiguousarray(p7, dtype=np.float64)
    df, db = _D(), _D()
    lib().or_weno_pair(ctypes.byref(df), ctypes.byref(db), _p(p), _D(ds))
    return df.value, db.value


def six_distance(v0, v1, v2, v3, v4, v5, s):
    f = lib().or_six_distance
    f.restype = _D
    return f(*[_D(x) for x in (v0, v1, v2, v3, v4, v5, s)])


def eno_quadratic_distance(v1, v0, v2, v3, ds):
    f = lib().or_eno_quadratic_distance
    f.restype = _D
    return f(*[_D(x) for x in (v1, v0, v2, v3, ds)])


def cubic_distance(v0, v1, v2, v3, s):
    f = lib().or_cubic_distance
    f.restype = _D
    return f(*[_D(x) for x in (v0, v1, v2, v3, s)])


def eno_cubic_pair(x7, h7):
    """(d_fore, d_back) of ENO_cubic_derivative on nodes x7 (x7[3] the point) with values h7"""
    x = np.ascontiguousarray(x7, dtype=np.float64)
    h = np.ascontiguousarray(h7, dtype=np.float64)
    df, db = _D(), _D()
    lib().or_eno_cubic_pair(ctypes.byref(df), ctypes.byref(db), _p(x), _p(h))
    return df.value, db.value


def wenork3_derivative_boundary(p7, r1, l1, ds, v_fore=0.0, v_back=0.0, rule=0):
    p = np.ascontiguousarray(p7, dtype=np.float64)
    df, db = _D(), _D()
    lib().or_wenork3_derivative_boundary(ctypes.byref(df), ctypes.byref(db), _p(p), _D(r1), _D(l1), _D(ds),
                                         _D(v_fore), _D(v_back), _I(rule))
    return df.value, db.value


def weno_derivative_boundary(p7, r1, l1, ds, v_fore=0.0, v_back=0.0, rule=0):
    p = np.ascontiguousarray(p7, dtype=np.float64)
    df, db = _D(), _D()
    lib().or_weno_derivative_boundary(ctypes.byref(df), ctypes.byref(db), _p(p), _D(r1), _D(l1), _D(ds),
                                      _D(v_fore), _D(v_back), _I(rule))
    return df.value, db.value


# --- wrappers ---------------------------------------------------------------

def set_distance(F, dx, dy, dz, order=6):
    rows, cols, pges = _dims(F)
    d6 = [np.empty_like(F) for _ in range(6)]
    md = np.empty_like(F)
    lib().or_set_distance(_I(order), _p(F), *[_p(a) for a in d6], _p(md), _I(rows), _I(cols), _I(pges),
                          _D(dx), _D(dy), _D(dz))
    return d6, md


def weno5_rk3_reinit(F, iterations, d6, min_dist, dx, dy, dz):
    F = f64(F).copy()
    rows, cols, pges = _dims(F)
    lib().or_weno5_rk3_reinit(_p(F), _I(iterations), *[_p(a) for a in d6], _p(min_dist), _I(rows), _I(cols),
                              _I(pges), _D(dx), _D(dy), _D(dz))
    return F


def eno2_reinit(F, iterations, dx, dy, dz, rk2=True, cfl=0.3):
    F = f64(F).copy()
    rows, cols, pges = _dims(F)
    lib().or_eno2_reinit(_p(F), _I(iterations), _I(1 if rk2 else 0), _D(cfl), _I(rows), _I(cols), _I(pges),
                         _D(dx), _D(dy), _D(dz))
    return F


def wenork3_reinit(F, iterations, dx, dy, dz):
    F = f64(F).copy()
    rows, cols, pges = _dims(F)
    lib().or_wenork3_reinit(_p(F), _I(iterations), _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz))
    return F


def variant_extend(C, iterations, F, dx, dy, dz, scheme):
    """ENORK2Extend ("enork2") or WENORK3Extend ("wenork3") of C off the zero set of F"""
    C = f64(C).copy()
    F = f64(F)
    rows, cols, pges = _dims(C)
    fn = {"enork2": lib().or_enork2_extend, "wenork3": lib().or_wenork3_extend}[scheme]
    fn(_p(C), _I(iterations), _p(F), _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz))
    return C


def extension_velocity(F, d6, floor, dx, dy, dz):
    rows, cols, pges = _dims(F)
    v = [np.empty_like(F) for _ in range(3)]
    lib().or_extension_velocity(_p(F), *[_p(a) for a in d6], _D(floor), *[_p(a) for a in v], _I(rows), _I(cols),
                                _I(pges), _D(dx), _D(dy), _D(dz))
    return v


def weno5_rk3_extend(C, iterations, F, d6, min_dist, dx, dy, dz):
    C = f64(C).copy()
    rows, cols, pges = _dims(C)
    lib().or_weno5_rk3_extend(_p(C), _I(iterations), _p(F), *[_p(a) for a in d6], _p(min_dist), _I(rows),
                              _I(cols), _I(pges), _D(dx), _D(dy), _D(dz))
    return C


def surface_redistance(A, F, iter1, iter2, dx, dy, dz):
    A = f64(A).copy()
    rows, cols, pges = _dims(A)
    lib().or_surface_redistance(_p(A), _p(F), _I(iter1), _I(iter2), _I(rows), _I(cols), _I(pges),
                                _D(dx), _D(dy), _D(dz))
    return A


def advect_rk3(F, steps, dt, v3, dx, dy, dz):
    F = f64(F).copy()
    rows, cols, pges = _dims(F)
    lib().or_advect_rk3(_p(F), _I(steps), _D(dt), *[_p(a) for a in v3], _I(rows), _I(cols), _I(pges),
                        _D(dx), _D(dy), _D(dz))
    return F


def area_volume_towers(F, dx, dy, dz):
    rows, cols, pges = _dims(F)
    a, v = _D(), _D()
    lib().or_area_volume_towers(_p(F), _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz),
                                ctypes.byref(a), ctypes.byref(v))
    return a.value, v.value


def area_volume_cosine(F, dx, dy, dz, eps):
    rows, cols, pges = _dims(F)
    a, v = _D(), _D()
    lib().or_area_volume_cosine(_p(F), _I(rows), _I(cols), _I(pges), _D(dx), _D(dy), _D(dz), _D(eps),
                                ctypes.byref(a), ctypes.byref(v))
    return a.value, v.value


def normal_velocity(F, v3, dx, dy, dz):
    rows, cols, pges = _dims(F)
    out = np.empty_like(F)
    lib().or_normal_velocity(_p(out), _p(F), *[_p(a) for a in v3], _I(rows), _I(cols), _I(pges),
                             _D(dx), _D(dy), _D(dz))
    return out


def pipeline(F0, psi0, u3, dt, eps, spacing, steps, ext_iters, reinit_every, reinit_iters, redist_iters,
             initial_reinit, first_step=0):
    """CPU restatement of the vesicle-step pipeline (same order as the product's
    pipeline.Pipeline.run).  Returns F, psi, Vn and the per-step (area, volume)."""
    dx, dy, dz = spacing
    d6, md = set_distance(F0, dx, dy, dz)
    F = weno5_rk3_reinit(F0, initial_reinit, d6, md, dx, dy, dz)
    psi = f64(psi0).copy()
    Vn = np.zeros_like(F)
    av = []
    for s in range(steps):
        i = first_step + s
        F = advect_rk3(F, 1, dt, u3, dx, dy, dz)
        psi = advect_rk3(psi, 1, dt, u3, dx, dy, dz)
        d6, md = set_distance(F, dx, dy, dz)
        Vn = normal_velocity(F, u3, dx, dy, dz)
        Vn = weno5_rk3_extend(Vn, ext_iters, F, d6, md, dx, dy, dz)
        if (i + 1) % reinit_every == 0:
            F = weno5_rk3_reinit(F, reinit_iters, d6, md, dx, dy, dz)
            psi = surface_redistance(psi, F, redist_iters[0], redist_iters[1], dx, dy, dz)
        av.append(area_volume_cosine(F, dx, dy, dz, eps))
    return F, psi, Vn, np.array(av)
