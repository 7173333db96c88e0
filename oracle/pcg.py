"""Python side of the Jacobi-PCG oracle (test infrastructure only); see pcg.c."""
import ctypes
import os

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", _HERE])
        _LIB = ctypes.CDLL(path)
        for name, real in (("oracle_pcg_f64", ctypes.c_double), ("oracle_pcg_f32", ctypes.c_float),
                           ("oracle_pcg_f32_d", ctypes.c_float)):
            fn = getattr(_LIB, name)
            P = ctypes.POINTER
            fn.argtypes = [ctypes.c_int, P(ctypes.c_int32), P(ctypes.c_int32), P(real), P(real), P(real),
                           ctypes.c_double, ctypes.c_int, ctypes.c_int, P(real), P(ctypes.c_double)]
            fn.restype = ctypes.c_int
    return _LIB


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def solve_pcg(a, b, tol=1e-5, max_iter=-1, res_fix=False, dtype=np.float64, dot_dtype=None):
    """solver.py:189-202 + solve_pcg.cu:15-136. Returns (x, iterations).

    dtype float64: all-FP64 oracle; float32: FP32 vectors and matrix, with FP32 dots
    (the reference, dot_dtype=None) or FP64 dots (dot_dtype=np.float64).
    """
    a = sp.csr_matrix(a)
    n = a.shape[0]
    if n == 0:
        return np.zeros(0, dtype), 0
    diag = a.diagonal()
    inv_diag = 1.0 / diag
    rowptr = a.indptr.astype(np.int32)
    col = a.indices.astype(np.int32)
    resid = np.zeros(1)
    if dtype == np.float64:
        ct, fn = ctypes.c_double, _lib().oracle_pcg_f64
    else:
        ct = ctypes.c_float
        fn = _lib().oracle_pcg_f32_d if dot_dtype == np.float64 else _lib().oracle_pcg_f32
    val = np.ascontiguousarray(a.data.astype(dtype))
    bb = np.ascontiguousarray(np.asarray(b).astype(dtype))
    idg = np.ascontiguousarray(inv_diag.astype(dtype))
    x = np.zeros(n, dtype)
    it = fn(n, _ptr(rowptr, ctypes.c_int32), _ptr(col, ctypes.c_int32), _ptr(val, ct), _ptr(bb, ct),
            _ptr(idg, ct), float(tol), int(max_iter), int(bool(res_fix)), _ptr(x, ct),
            resid.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return x, it


def relative_residual(a, x, b):
    """||b - A x|| / ||b|| in FP64."""
    a = sp.csr_matrix(a).astype(np.float64)
    r = np.asarray(b, np.float64) - a @ np.asarray(x, np.float64)
    nb = np.linalg.norm(np.asarray(b, np.float64))
    return float(np.linalg.norm(r) / nb) if nb > 0 else float(np.linalg.norm(r))


class Prepared:
    """CSR arrays converted once (float32 values, FP64 dots) for repeated timed runs."""

    def __init__(self, a=None, arrays=None):
        if arrays is not None:
            rowptr, col, val, inv_diag = arrays
            self.n = inv_diag.shape[0]
            self.rowptr = np.ascontiguousarray(rowptr, np.int32)
            self.col = np.ascontiguousarray(col, np.int32)
            self.val = np.ascontiguousarray(val, np.float32)
            self.inv_diag = np.ascontiguousarray(inv_diag, np.float32)
            return
        a = sp.csr_matrix(a)
        self.n = a.shape[0]
        self.rowptr = a.indptr.astype(np.int32)
        self.col = a.indices.astype(np.int32)
        self.val = np.ascontiguousarray(a.data.astype(np.float32))
        self.inv_diag = np.ascontiguousarray((1.0 / a.diagonal()).astype(np.float32))

    def run(self, b, iters, tol=0.0):
        x = np.zeros(self.n, np.float32)
        bb = np.ascontiguousarray(np.asarray(b, np.float32))
        resid = np.zeros(1)
        ct = ctypes.c_float
        it = _lib().oracle_pcg_f32_d(self.n, _ptr(self.rowptr, ctypes.c_int32), _ptr(self.col, ctypes.c_int32),
                                     _ptr(self.val, ct), _ptr(bb, ct), _ptr(self.inv_diag, ct), float(tol), int(iters), 0,
                                     _ptr(x, ct), resid.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return x, it
