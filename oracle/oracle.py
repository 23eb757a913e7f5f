"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the gb_amd package.  Each function
names the gb v0.16 routine its C counterpart restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from gb_amd.matrix import F4Matrix, SparseRows

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


class _Mat(C.Structure):
    _fields_ = [("nr", C.c_int32), ("nc", C.c_int32), ("ncl", C.c_int32),
                ("rp", C.c_void_p), ("ci", C.c_void_p), ("rcf", C.c_void_p),
                ("cp", C.c_void_p), ("cf", C.c_void_p)]


class _Rows(C.Structure):
    _fields_ = [("np", C.c_int32), ("rp", C.POINTER(C.c_int64)),
                ("ci", C.POINTER(C.c_int32)), ("cf", C.POINTER(C.c_uint32))]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_arith_bits.argtypes = [C.c_int32]
        L.orc_arith_bits.restype = C.c_int32
        L.orc_mod_inverse.argtypes = [C.c_int32, C.c_int32]
        L.orc_mod_inverse.restype = C.c_int32
        L.orc_normalize_dense_row.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.orc_sparse_pass.argtypes = [C.POINTER(_Mat), C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
        L.orc_sparse_pass.restype = C.c_int64
        L.orc_sparse_pass_work.argtypes = [C.POINTER(_Mat), C.c_int32, C.c_int32,
                                           C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_sparse_pass_work.restype = C.c_int64
        L.orc_dense_rref.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_int32,
                                     C.c_void_p, C.c_void_p]
        L.orc_dense_rref.restype = C.c_int32
        for fn in (L.orc_la_sparse_dense, L.orc_la_sparse):
            fn.argtypes = [C.POINTER(_Mat), C.c_int32, C.c_int32, C.POINTER(_Rows)]
            fn.restype = C.c_int32
        L.orc_interreduce.argtypes = [C.POINTER(_Mat), C.c_int32, C.POINTER(_Rows)]
        L.orc_interreduce.restype = C.c_int32
        L.orc_rows_free.argtypes = [C.POINTER(_Rows)]
        _lib = L
    return _lib


def _mat(m: F4Matrix):
    arrs = (np.ascontiguousarray(m.rp, np.int64), np.ascontiguousarray(m.ci, np.int32),
            np.ascontiguousarray(m.rcf, np.int32), np.ascontiguousarray(m.cp, np.int64),
            np.ascontiguousarray(m.cf, np.uint32))
    s = _Mat(m.nr, m.nc, m.ncl, *(a.ctypes.data for a in arrs))
    return s, arrs


def _take_rows(r: _Rows) -> SparseRows:
    n = r.np
    rp = np.ctypeslib.as_array(r.rp, shape=(n + 1,)).copy()
    nnz = int(rp[-1])
    ci = np.ctypeslib.as_array(r.ci, shape=(max(nnz, 1),))[:nnz].copy()
    cf = np.ctypeslib.as_array(r.cf, shape=(max(nnz, 1),))[:nnz].copy()
    lib().orc_rows_free(C.byref(r))
    return SparseRows(rp=rp, ci=ci, cf=cf)


def arith_bits(p: int) -> int:
    """io.c:815-829 and io.c:710-729."""
    return lib().orc_arith_bits(p)


def mod_inverse(a: int, p: int) -> int:
    """tools.c:83-110."""
    return lib().orc_mod_inverse(a, p)


def normalize_dense_row(row, p: int) -> np.ndarray:
    """la_ff_32.c:24-60 on an int32 row (entries need not be reduced)."""
    r = np.ascontiguousarray(row, dtype=np.int32).copy()
    lib().orc_normalize_dense_row(r.ctypes.data, len(r), p)
    return r


def sparse_pass(m: F4Matrix, nthreads: int = 1):
    """la_ff_32.c:1081-1179: the reduced D' rows (dense, ncr wide) in the
    reference's order, and the input row each one came from."""
    s, keep = _mat(m)
    nrl = m.nr
    out = np.zeros((max(nrl, 1), max(m.ncr, 1)), dtype=np.uint32)
    src = np.zeros(max(nrl, 1), dtype=np.int32)
    n = lib().orc_sparse_pass(C.byref(s), m.p, nthreads, out.ctypes.data, src.ctypes.data)
    if n < 0:
        raise ValueError("matrix violates the known-pivot contract")
    return out[:n, :m.ncr], src[:n]


def sparse_pass_work(m: F4Matrix, nthreads: int = 1):
    """(#nonzero D' rows, pivots applied, pivot entries streamed)."""
    s, keep = _mat(m)
    a = C.c_int64(0)
    e = C.c_int64(0)
    n = lib().orc_sparse_pass_work(C.byref(s), m.p, nthreads, C.byref(a), C.byref(e))
    return int(n), int(a.value), int(e.value)


def dense_rref(dm: np.ndarray, p: int, nthreads: int = 1):
    """la_ff_32.c:1217-1333, 1181-1215: (RREF rows in decreasing lead, leads)."""
    dm = np.ascontiguousarray(dm, dtype=np.uint32)
    m, n = dm.shape
    out = np.zeros((max(min(m, n), 1), max(n, 1)), dtype=np.uint32)
    lead = np.zeros(max(min(m, n), 1), dtype=np.int32)
    r = lib().orc_dense_rref(m, n, dm.ctypes.data, p, nthreads, out.ctypes.data, lead.ctypes.data)
    return out[:r, :n], lead[:r]


def la_sparse_dense(m: F4Matrix, nthreads: int = 1) -> SparseRows:
    """la_option 1: la_ff_32.c:1804-1851."""
    s, keep = _mat(m)
    r = _Rows()
    if lib().orc_la_sparse_dense(C.byref(s), m.p, nthreads, C.byref(r)) < 0:
        raise ValueError("matrix violates the known-pivot contract")
    return _take_rows(r)


def la_sparse(m: F4Matrix, nthreads: int = 1) -> SparseRows:
    """la_option 2: la_ff_32.c:1774-1802, 945-1079."""
    s, keep = _mat(m)
    r = _Rows()
    if lib().orc_la_sparse(C.byref(s), m.p, nthreads, C.byref(r)) < 0:
        raise ValueError("row lead outside the matrix")
    return _take_rows(r)


def interreduce(m: F4Matrix) -> SparseRows:
    """la_ff_32.c:1949-2017 (ncl = 0, ascending lead output)."""
    s, keep = _mat(m)
    r = _Rows()
    if lib().orc_interreduce(C.byref(s), m.p, C.byref(r)) < 0:
        raise ValueError("row lead outside the matrix")
    return _take_rows(r)


def f4(system, nthreads: int = 1, reduce_gb: int = 0, info_level: int = 0):
    """The F4 host (libgbgpu's gbgpu_f4) with the oracle's linear algebra
    in every round (orc_f4_la): an end-to-end run on the CPU."""
    from gb_amd import native
    L = lib()
    fn = C.cast(L.orc_f4_la, C.c_void_p).value
    nth = C.c_int(nthreads)
    return native.f4(system, fn, C.cast(C.pointer(nth), C.c_void_p), reduce_gb, info_level)
