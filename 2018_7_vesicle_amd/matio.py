"""MATLAB Level 5 MAT-files through the library's reader/writer
(vls_mat_* of include/vesicle_ls.h): the reference's fixtures and run records
(src/7_0_paper/2_0_Vesicle/pear.mat, CFL*.mat, oblate*.mat) and restart files.

Arrays come back as numpy float64 arrays of MATLAB's shape in MATLAB's
(column-major) order; `field` turns a 3-D [rows, cols, pges] array into the
(pges, cols, rows) layout of the level-set fields.
"""
import ctypes

import numpy as np

from ._lib import VLSError, call


def _b(s):
    return s.encode() if isinstance(s, str) else s


def whosmat(path):
    """[(name, shape)] of the numeric arrays in the file"""
    n = ctypes.c_int()
    call("vls_mat_count", _b(str(path)), ctypes.byref(n))
    out = []
    for k in range(n.value):
        name = ctypes.create_string_buffer(64)
        nd = ctypes.c_int()
        dims = (ctypes.c_int64 * 8)()
        call("vls_mat_info", _b(str(path)), k, name, 64, ctypes.byref(nd), dims, None)
        out.append((name.value.decode(), tuple(int(dims[i]) for i in range(nd.value))))
    return out


def loadmat(path):
    """{name: array} of every numeric array (real part, as float64)"""
    out = {}
    for name, shape in whosmat(path):
        buf = np.empty(int(np.prod(shape)), dtype=np.float64)
        call("vls_mat_read", _b(str(path)), _b(name), buf.ctypes.data_as(ctypes.c_void_p), buf.size)
        out[name] = buf.reshape(shape, order="F")
    return out


def savemat(path, arrays, compress=True):
    """write {name: array} as double arrays (vectors become 1 x n rows, as MATLAB's)"""
    names, keep, ndims, dims = [], [], [], []
    for name, a in arrays.items():
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(1, -1)
        if a.ndim > 8:
            raise VLSError("MAT arrays have at most 8 dimensions")
        names.append(_b(name))
        keep.append(np.asfortranarray(a))
        ndims.append(a.ndim)
        dims.extend(list(a.shape) + [0] * (8 - a.ndim))
    n = len(names)
    c_names = (ctypes.c_char_p * max(n, 1))(*names)
    c_data = (ctypes.c_void_p * max(n, 1))(*[k.ctypes.data for k in keep])
    c_nd = (ctypes.c_int * max(n, 1))(*ndims)
    c_dims = (ctypes.c_int64 * max(8 * n, 1))(*dims)
    call("vls_mat_write", _b(str(path)), n, c_names, c_data, c_nd, c_dims, int(bool(compress)))


def field(a):
    """a MATLAB [rows, cols, pges] array as a C-contiguous (pges, cols, rows) field"""
    if a.ndim != 3:
        raise VLSError("a level-set field is a 3-D array")
    return np.ascontiguousarray(a.transpose(2, 1, 0))


def matlab_array(f):
    """the inverse of field: a (pges, cols, rows) field as a MATLAB [rows, cols, pges] array"""
    return np.asfortranarray(np.asarray(f).transpose(2, 1, 0))
