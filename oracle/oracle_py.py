"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU parity oracle
(oracle/build/liboracle.so).  Imported by tests/, the smoke check and the
CPU-baseline leg of bench.py; never by the product package."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


def build(force: bool = False):
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(st, what):
    if st != 0:
        raise RuntimeError(f"oracle {what} failed with status {st}")


def residual(b, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    res = np.zeros_like(q)
    _check(lib().oracle_residual(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys), _d(q), _d(res)),
           "residual")
    return res


def residual_status(b, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    res = np.zeros_like(q)
    return lib().oracle_residual(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys), _d(q), _d(res))


def residual_scale(b, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    res = np.zeros_like(q)
    _check(lib().oracle_residual_scale(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys), _d(q), _d(res)),
           "residual_scale")
    return res


def residual_complex(b, qc):
    qr = np.ascontiguousarray(qc.real, dtype=np.float64)
    qi = np.ascontiguousarray(qc.imag, dtype=np.float64)
    rr = np.zeros_like(qr)
    ri = np.zeros_like(qr)
    _check(lib().oracle_residual_complex(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys),
                                         _d(qr), _d(qi), _d(rr), _d(ri)), "residual_complex")
    return rr + 1j * ri


def jv_complex_step(b, q, v, h=1e-20):
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros_like(q)
    _check(lib().oracle_jv_complex_step(C.byref(b.mesh), C.byref(b.sbp), C.byref(b.phys),
                                        _d(q), _d(v), C.c_double(h), _d(out)), "jv")
    return out


def sparsity_bsr(b):
    nrows = b.mesh.numEl * b.sbp.Np
    rowptr = np.zeros(nrows + 1, dtype=np.int64)
    L = lib()
    L.oracle_sparsity_bsr.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    _check(L.oracle_sparsity_bsr(C.addressof(b.mesh), C.addressof(b.sbp),
                                 rowptr.ctypes.data, None), "sparsity")
    colind = np.zeros(rowptr[-1], dtype=np.int32)
    _check(L.oracle_sparsity_bsr(C.addressof(b.mesh), C.addressof(b.sbp),
                                 rowptr.ctypes.data, colind.ctypes.data), "sparsity")
    return rowptr, colind


def color_elements(b):
    color = np.zeros(b.mesh.numEl, dtype=np.int32)
    nc = C.c_int32(0)
    _check(lib().oracle_color_elements(C.byref(b.mesh), color.ctypes.data_as(C.POINTER(C.c_int32)),
                                       C.byref(nc)), "coloring")
    return color, nc.value


def jacobian_cs(b, q, rowptr, colind):
    nd = b.mesh.dim + 2
    q = np.ascontiguousarray(q, dtype=np.float64)
    val = np.zeros((colind.shape[0], nd, nd))
    L = lib()
    L.oracle_jacobian_cs.argtypes = [C.c_void_p] * 7
    _check(L.oracle_jacobian_cs(C.addressof(b.mesh), C.addressof(b.sbp), C.addressof(b.phys),
                                q.ctypes.data, rowptr.ctypes.data, colind.ctypes.data,
                                val.ctypes.data), "jacobian_cs")
    return val
