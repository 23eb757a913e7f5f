"""ctypes binding of libgbgpu.so (the C-ABI declared in include/gbgpu.h).

There is no CPU fallback: every entry point runs the HIP kernels, and
loading fails loudly when the library or the GPU is missing.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
import os
import re

import numpy as np

from .matrix import F4Matrix, SparseRows

HERE = os.path.dirname(os.path.abspath(__file__))
# GBGPU_LIB may point at another build of the same C-ABI
LIB_PATH = os.environ.get("GBGPU_LIB", os.path.join(HERE, "libgbgpu.so"))
HEADERS = [os.path.join(os.path.dirname(HERE), "include", h)
           for h in ("gbgpu.h", "gbgpu_gb.h", "gbgpu_f4.h")]

FORM_PLAIN = 0
FORM_MFMA = 1

_lib = None


class GpuError(RuntimeError):
    pass


class _Csr(C.Structure):
    _fields_ = [("nr", C.c_int32), ("nc", C.c_int32), ("ncl", C.c_int32), ("ncf", C.c_int32),
                ("p", C.c_uint32), ("rp", C.c_void_p), ("ci", C.c_void_p), ("rcf", C.c_void_p),
                ("cp", C.c_void_p), ("cf", C.c_void_p)]


class _Rows(C.Structure):
    _fields_ = [("np", C.c_int32), ("rp", C.POINTER(C.c_int64)),
                ("ci", C.POINTER(C.c_int32)), ("cf", C.POINTER(C.c_uint32))]


class _F4In(C.Structure):
    _fields_ = [("nr_vars", C.c_int32), ("nr_gens", C.c_int32),
                ("lens", C.POINTER(C.c_int32)), ("exps", C.POINTER(C.c_int32)),
                ("cfs", C.POINTER(C.c_int32)), ("field_char", C.c_int32),
                ("mon_order", C.c_int32), ("la_option", C.c_int32), ("reduce_gb", C.c_int32),
                ("info_level", C.c_int32)]


class _F4Out(C.Structure):
    _fields_ = [("bld", C.c_int32), ("blen", C.POINTER(C.c_int32)),
                ("bexp", C.POINTER(C.c_int32)), ("bcf", C.POINTER(C.c_uint32)),
                ("nterms", C.c_int64), ("rounds", C.c_int32), ("max_rows", C.c_int64),
                ("max_cols", C.c_int64), ("max_nnz", C.c_int64), ("rows_total", C.c_int64),
                ("la_ms", C.c_double), ("total_ms", C.c_double), ("select_ms", C.c_double),
                ("symbolic_ms", C.c_double), ("build_ms", C.c_double), ("update_ms", C.c_double),
                ("trace", C.POINTER(C.c_int64))]

# fields of one round's record in _F4Out.trace (GBGPU_F4_TRACE_W in gbgpu_f4.h)
F4_TRACE_FIELDS = ("degree", "pairs", "rows", "cols", "pivot_cols", "nnz", "new", "la_hash")


class _F4MultiStats(C.Structure):
    _fields_ = [("batches", C.c_int64), ("matrices", C.c_int64), ("device_ms", C.c_double)]


class _F4Gpu(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("form", C.c_int32), ("device_ms", C.c_double)]


@dataclass
class PolySystem:
    """A polynomial system in f4_julia's arrays: lens[i] terms per
    generator, exps nv exponents per term, cfs one coefficient per term."""
    nv: int
    p: int
    lens: np.ndarray
    exps: np.ndarray
    cfs: np.ndarray
    order: int = 0  # 0 DRL, 1 LEX


@dataclass
class Basis:
    """A basis returned by the F4 host, with the run's statistics."""
    nv: int
    lens: np.ndarray
    exps: np.ndarray  # nterms x nv
    cfs: np.ndarray
    stats: dict

    def polys(self):
        """Elements as lists of (exponent tuple, coefficient), decreasing."""
        out, t = [], 0
        for L in self.lens:
            out.append([(tuple(int(x) for x in self.exps[t + k]), int(self.cfs[t + k]))
                        for k in range(int(L))])
            t += int(L)
        return out

    def leads(self):
        return sorted(p[0][0] for p in self.polys())


def f4(system: PolySystem, la_fn: int, la_user, reduce_gb: int = 0, info_level: int = 0,
       keep=None) -> Basis:
    """Runs the F4 host with the linear-algebra callback at address la_fn
    (a gbgpu_f4_la_t) and its user pointer."""
    L = lib()
    inp, _arrays = _f4_in(system, reduce_gb, info_level)
    out = _F4Out()
    rc = L.gbgpu_f4(C.byref(inp), C.c_void_p(la_fn), la_user, C.byref(out))
    if rc != 0:
        raise GpuError(f"gbgpu_f4 failed ({rc})")
    b = _basis(out, system.nv)
    L.gbgpu_f4_out_free(C.byref(out))
    return b


def _f4_in(system: PolySystem, reduce_gb: int, info_level: int):
    lens = np.ascontiguousarray(system.lens, np.int32)
    exps = np.ascontiguousarray(system.exps, np.int32).reshape(-1)
    cfs = np.ascontiguousarray(system.cfs, np.int32)
    pi = C.POINTER(C.c_int32)
    inp = _F4In(system.nv, len(lens), lens.ctypes.data_as(pi), exps.ctypes.data_as(pi),
                cfs.ctypes.data_as(pi), system.p, system.order, 1, reduce_gb, info_level)
    return inp, (lens, exps, cfs)


def _basis(out: "_F4Out", nv: int) -> Basis:
    nt = int(out.nterms)
    system_nv = nv
    return Basis(nv=system_nv,
              lens=np.ctypeslib.as_array(out.blen, shape=(out.bld,)).copy() if out.bld else np.zeros(0, np.int32),
              exps=np.ctypeslib.as_array(out.bexp, shape=(nt * nv,)).copy().reshape(nt, nv) if nt else np.zeros((0, nv), np.int32),
              cfs=np.ctypeslib.as_array(out.bcf, shape=(nt,)).copy() if nt else np.zeros(0, np.uint32),
              stats={"rounds": out.rounds, "max_rows": out.max_rows, "max_cols": out.max_cols,
                     "max_nnz": out.max_nnz, "rows_total": out.rows_total, "la_ms": out.la_ms,
                     "total_ms": out.total_ms, "select_ms": out.select_ms,
                     "symbolic_ms": out.symbolic_ms, "build_ms": out.build_ms,
                     "update_ms": out.update_ms,
                     "trace": ([dict(zip(F4_TRACE_FIELDS, map(int, row))) for row in
                                np.ctypeslib.as_array(out.trace, shape=(out.rounds, len(F4_TRACE_FIELDS)))]
                               if out.rounds and out.trace else [])})


def declared_symbols() -> list[str]:
    """Every function declared in include/gbgpu.h and include/gbgpu_gb.h."""
    names = set()
    for h in HEADERS:
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names.update(re.findall(r"\b(gbgpu_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def lib():
    """Load libgbgpu.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GpuError(f"{LIB_PATH} is missing: build it with `python -m gb_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, u32, i64 = C.c_void_p, C.c_int32, C.c_uint32, C.c_int64
    pi32 = C.POINTER(C.c_int32)
    sig = {
        "gbgpu_init": ([C.c_int, C.POINTER(vp)], C.c_int),
        "gbgpu_finalize": ([vp], None),
        "gbgpu_set_stream": ([vp, vp], C.c_int),
        "gbgpu_last_error": ([vp], C.c_char_p),
        "gbgpu_arith_bits": ([u32], C.c_int),
        "gbgpu_phase_ms": ([vp, C.c_int], C.c_double),
        "gbgpu_mat_upload": ([vp, C.POINTER(_Csr), C.POINTER(vp)], C.c_int),
        "gbgpu_mat_free": ([vp], None),
        "gbgpu_mat_upload_rows": ([vp, C.POINTER(_Csr), vp, i64, vp, vp, i32, vp, C.POINTER(vp)],
                                  C.c_int),
        "gbgpu_cfstore_create": ([vp, u32, C.POINTER(vp)], C.c_int),
        "gbgpu_cfstore_append": ([vp, vp, i64, C.POINTER(C.c_int64)], C.c_int),
        "gbgpu_cfstore_size": ([vp], C.c_int64),
        "gbgpu_cfstore_free": ([vp], None),
        "gbgpu_sparse_pass": ([vp, vp, C.c_int, pi32], C.c_int),
        "gbgpu_sparse_pass_download": ([vp, vp, vp, vp], C.c_int),
        "gbgpu_sparse_pass_work": ([vp, vp, pi32, C.POINTER(C.c_uint64)], C.c_int),
        "gbgpu_la_sparse_dense": ([vp, vp, C.c_int, pi32], C.c_int),
        "gbgpu_la_sparse": ([vp, vp, C.c_int, pi32], C.c_int),
        "gbgpu_interreduce": ([vp, vp, C.c_int, pi32], C.c_int),
        "gbgpu_rows_download": ([vp, vp, C.POINTER(_Rows)], C.c_int),
        "gbgpu_rows_free": ([C.POINTER(_Rows)], None),
        "gbgpu_la_ff": ([vp, C.POINTER(_Csr), C.c_int, C.c_int, C.POINTER(_Rows)], C.c_int),
        "gbgpu_dense_upload": ([vp, i32, i32, u32, vp, C.POINTER(vp)], C.c_int),
        "gbgpu_dense_free": ([vp], None),
        "gbgpu_dense_rref": ([vp, vp, C.c_int, pi32], C.c_int),
        "gbgpu_dense_download": ([vp, vp, vp, vp], C.c_int),
        "gbgpu_dense_rref_noalloc_bytes": ([i32, i32, u32, C.c_int], C.c_size_t),
        "gbgpu_dense_rref_noalloc": ([vp, vp, i32, i32, u32, C.c_int, vp, vp, pi32], C.c_int),
        "gbgpu_batch_upload": ([vp, C.POINTER(_Csr), i32, C.POINTER(vp)], C.c_int),
        "gbgpu_batch_la": ([vp, vp, C.c_int, pi32], C.c_int),
        "gbgpu_batch_rows_download": ([vp, vp, i32, C.POINTER(_Rows)], C.c_int),
        "gbgpu_batch_info": ([vp, pi32, i32], C.c_int),
        "gbgpu_batch_free": ([vp], None),
        "gbgpu_f4": ([C.POINTER(_F4In), vp, vp, C.POINTER(_F4Out)], C.c_int),
        "gbgpu_f4_multimod": ([vp, C.POINTER(_F4In), i32, i32, C.POINTER(_F4Out),
                               C.POINTER(_F4MultiStats)], C.c_int),
        "gbgpu_f4_out_free": ([C.POINTER(_F4Out)], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _ = i64
    _lib = L
    return L


def lib_checked_gpu() -> "Context":
    """A context on device 0; raises when no GPU is usable."""
    return Context(0)


def _csr(m: F4Matrix):
    arrs = (np.ascontiguousarray(m.rp, np.int64), np.ascontiguousarray(m.ci, np.int32),
            np.ascontiguousarray(m.rcf, np.int32), np.ascontiguousarray(m.cp, np.int64),
            np.ascontiguousarray(m.cf, np.uint32))
    s = _Csr(m.nr, m.nc, m.ncl, len(m.cp) - 1, m.p, *(a.ctypes.data for a in arrs))
    return s, arrs


def _take_rows(r: _Rows) -> SparseRows:
    n = r.np
    rp = np.ctypeslib.as_array(r.rp, shape=(n + 1,)).copy()
    nnz = int(rp[-1])
    if nnz:
        ci = np.ctypeslib.as_array(r.ci, shape=(nnz,)).copy()
        cf = np.ctypeslib.as_array(r.cf, shape=(nnz,)).copy()
    else:
        ci = np.zeros(0, np.int32)
        cf = np.zeros(0, np.uint32)
    lib().gbgpu_rows_free(C.byref(r))
    return SparseRows(rp=rp, ci=ci, cf=cf)


class Context:
    """One device, one stream (gbgpu_ctx)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        rc = L.gbgpu_init(device, C.byref(h))
        if rc != 0:
            raise GpuError(f"gbgpu_init({device}) failed with {rc}: no usable GPU")
        self.h = h

    def close(self):
        if self.h:
            lib().gbgpu_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = lib().gbgpu_last_error(self.h).decode()
            raise GpuError(f"{what} failed ({rc}): {msg}")

    def set_stream(self, stream_ptr: int):
        self._check(lib().gbgpu_set_stream(self.h, C.c_void_p(stream_ptr)), "set_stream")

    def phase_ms(self, phase: int) -> float:
        return float(lib().gbgpu_phase_ms(self.h, phase))

    def upload(self, m: F4Matrix) -> "DeviceMatrix":
        s, keep = _csr(m)
        h = C.c_void_p()
        self._check(lib().gbgpu_mat_upload(self.h, C.byref(s), C.byref(h)), "upload")
        return DeviceMatrix(self, h, m)

    def upload_rows(self, m: F4Matrix, mon: np.ndarray, src: np.ndarray,
                    colmap: np.ndarray, store: "CfStore | None" = None,
                    cp: np.ndarray | None = None) -> "DeviceMatrix":
        """Upload m with its column indices formed on the device
        (gbgpu_mat_upload_rows): row q is mon[src[q] : src[q] + len(q)]
        mapped through colmap; m.ci is not sent.  With a store, the
        coefficient arrays are the store's at offsets cp (m.cp, m.cf unused)."""
        s, keep = _csr(m)
        s.ci = None
        if store is not None:
            cp = np.ascontiguousarray(cp, np.int64)
            s.ncf = len(cp) - 1
            s.cp = cp.ctypes.data
            s.cf = None
        mon = np.ascontiguousarray(mon, np.int32)
        src = np.ascontiguousarray(src, np.int64)
        colmap = np.ascontiguousarray(colmap, np.int32)
        h = C.c_void_p()
        self._check(lib().gbgpu_mat_upload_rows(self.h, C.byref(s), mon.ctypes.data, len(mon),
                                                src.ctypes.data, colmap.ctypes.data, len(colmap),
                                                store.h if store is not None else None,
                                                C.byref(h)), "upload_rows")
        return DeviceMatrix(self, h, m)

    def cfstore(self, p: int) -> "CfStore":
        h = C.c_void_p()
        self._check(lib().gbgpu_cfstore_create(self.h, p, C.byref(h)), "cfstore")
        return CfStore(self, h)

    def la_ff(self, m: F4Matrix, la_option: int = 1, form: int = FORM_PLAIN) -> SparseRows:
        s, keep = _csr(m)
        r = _Rows()
        self._check(lib().gbgpu_la_ff(self.h, C.byref(s), la_option, form, C.byref(r)), "la_ff")
        return _take_rows(r)

    def batch(self, mats) -> "DeviceBatch":
        """Upload many F4 matrices (each with its own prime) as one batch."""
        mats = list(mats)
        structs, keep = [], []
        for m in mats:
            st, k = _csr(m)
            structs.append(st)
            keep.append(k)
        arr = (_Csr * max(len(structs), 1))(*structs)
        h = C.c_void_p()
        self._check(lib().gbgpu_batch_upload(self.h, arr, len(structs), C.byref(h)),
                    "batch upload")
        return DeviceBatch(self, h, len(mats))

    def f4(self, system: PolySystem, form: int = FORM_MFMA, reduce_gb: int = 0,
           info_level: int = 0) -> Basis:
        """F4 on this device: every round's linear algebra on the GPU."""
        L = lib()
        g = _F4Gpu(self.h, form, 0.0)
        fn = C.cast(L.gbgpu_f4_gpu_la, C.c_void_p).value
        b = f4(system, fn, C.cast(C.pointer(g), C.c_void_p), reduce_gb, info_level)
        b.stats["device_ms"] = g.device_ms
        return b

    def f4_multimod(self, systems, form: int = FORM_MFMA, reduce_gb: int = 0):
        """F4 of several images (one PolySystem per prime) in lockstep, every
        round's matrices of all images in one batched launch
        (gbgpu_f4_multimod).  Returns (bases, stats)."""
        L = lib()
        n = len(systems)
        ins = (_F4In * max(n, 1))()
        keep = []
        for i, s in enumerate(systems):
            ins[i], arrs = _f4_in(s, reduce_gb, 0)
            keep.append(arrs)
        outs = (_F4Out * max(n, 1))()
        st = _F4MultiStats()
        rc = L.gbgpu_f4_multimod(self.h, ins, n, form, outs, C.byref(st))
        try:
            if rc != 0:
                raise GpuError(f"gbgpu_f4_multimod failed ({rc}): {L.gbgpu_last_error(self.h).decode()}")
            bases = [_basis(outs[i], systems[i].nv) for i in range(n)]
        finally:
            for i in range(n):
                L.gbgpu_f4_out_free(C.byref(outs[i]))
        return bases, {"batches": int(st.batches), "matrices": int(st.matrices),
                       "device_ms": float(st.device_ms)}

    def dense_rref_noalloc(self, a_dev: int, m: int, n: int, p: int, form: int, rows_dev: int,
                           lead_dev: int) -> int:
        """gbgpu_dense_rref_noalloc on device buffers the caller owns (raw
        device addresses): a (m x n, overwritten), rows (min(m, n) x n) and
        lead (min(m, n)); returns the rank."""
        r = C.c_int32(0)
        self._check(lib().gbgpu_dense_rref_noalloc(self.h, C.c_void_p(a_dev), m, n, p, form,
                                                   C.c_void_p(rows_dev), C.c_void_p(lead_dev),
                                                   C.byref(r)), "dense rref (no allocation)")
        return int(r.value)

    def dense(self, a: np.ndarray, p: int) -> "DeviceDense":
        a = np.ascontiguousarray(a, dtype=np.uint32)
        h = C.c_void_p()
        self._check(lib().gbgpu_dense_upload(self.h, a.shape[0], a.shape[1], p,
                                             a.ctypes.data, C.byref(h)), "dense upload")
        return DeviceDense(self, h, a.shape, p)


class CfStore:
    """Coefficient arrays kept on the device (gbgpu_cfstore)."""

    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def append(self, cf: np.ndarray) -> int:
        cf = np.ascontiguousarray(cf, np.uint32)
        off = C.c_int64()
        self.ctx._check(lib().gbgpu_cfstore_append(self.h, cf.ctypes.data, len(cf), C.byref(off)),
                        "cfstore append")
        return int(off.value)

    def size(self) -> int:
        return int(lib().gbgpu_cfstore_size(self.h))

    def free(self):
        if self.h:
            lib().gbgpu_cfstore_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceMatrix:
    """An F4 matrix resident on the device (gbgpu_mat)."""

    def __init__(self, ctx: Context, h, m: F4Matrix):
        self.ctx, self.h, self.m = ctx, h, m

    def free(self):
        if self.h:
            lib().gbgpu_mat_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def sparse_pass(self, form: int = FORM_PLAIN) -> int:
        n = C.c_int32(0)
        self.ctx._check(lib().gbgpu_sparse_pass(self.ctx.h, self.h, form, C.byref(n)), "sparse pass")
        return int(n.value)

    def sparse_pass_work(self):
        """(#nonzero D' rows, pivots applied, pivot entries streamed)."""
        n = C.c_int32(0)
        w = (C.c_uint64 * 2)()
        self.ctx._check(lib().gbgpu_sparse_pass_work(self.ctx.h, self.h, C.byref(n), w), "work")
        return int(n.value), int(w[0]), int(w[1])

    def sparse_pass_result(self, nd: int):
        ncr = self.m.ncr
        dense = np.zeros((max(nd, 1), max(ncr, 1)), dtype=np.uint32)
        src = np.zeros(max(nd, 1), dtype=np.int32)
        self.ctx._check(lib().gbgpu_sparse_pass_download(self.ctx.h, self.h, dense.ctypes.data,
                                                         src.ctypes.data), "download")
        return dense[:nd, :ncr], src[:nd]

    def la_sparse_dense(self, form: int = FORM_PLAIN) -> int:
        n = C.c_int32(0)
        self.ctx._check(lib().gbgpu_la_sparse_dense(self.ctx.h, self.h, form, C.byref(n)), "la 1")
        return int(n.value)

    def la_sparse(self, form: int = FORM_PLAIN) -> int:
        n = C.c_int32(0)
        self.ctx._check(lib().gbgpu_la_sparse(self.ctx.h, self.h, form, C.byref(n)), "la 2")
        return int(n.value)

    def interreduce(self, form: int = FORM_PLAIN) -> int:
        n = C.c_int32(0)
        self.ctx._check(lib().gbgpu_interreduce(self.ctx.h, self.h, form, C.byref(n)), "interreduce")
        return int(n.value)

    def rows(self) -> SparseRows:
        r = _Rows()
        self.ctx._check(lib().gbgpu_rows_download(self.ctx.h, self.h, C.byref(r)), "rows download")
        return _take_rows(r)


class DeviceBatch:
    """Many F4 matrices resident on the device (gbgpu_batch)."""

    def __init__(self, ctx: Context, h, nmat: int):
        self.ctx, self.h, self.nmat = ctx, h, nmat

    def free(self):
        if self.h:
            lib().gbgpu_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def la(self, form: int = FORM_MFMA) -> np.ndarray:
        """la_option 1 on every matrix (one graph launch); new pivots per matrix."""
        npv = np.zeros(max(self.nmat, 1), dtype=np.int32)
        self.ctx._check(lib().gbgpu_batch_la(self.ctx.h, self.h, form,
                                             npv.ctypes.data_as(C.POINTER(C.c_int32))),
                        "batch la")
        return npv[: self.nmat]

    def rows(self, i: int) -> SparseRows:
        r = _Rows()
        self.ctx._check(lib().gbgpu_batch_rows_download(self.ctx.h, self.h, i, C.byref(r)),
                        "batch rows download")
        return _take_rows(r)

    def info(self) -> dict:
        v = np.zeros(8, dtype=np.int32)
        lib().gbgpu_batch_info(self.h, v.ctypes.data_as(C.POINTER(C.c_int32)), 8)
        return {"matrices": int(v[0]), "window_shift": int(v[1]), "waves_per_cu": int(v[2]),
                "sparse_grid": int(v[3]), "replays": int(v[4]), "graph_nodes": int(v[5]),
                "dense_tile": [int(v[6]), int(v[7])]}


class DeviceDense:
    """A dense block resident on the device (gbgpu_dense)."""

    def __init__(self, ctx: Context, h, shape, p: int):
        self.ctx, self.h, self.shape, self.p = ctx, h, shape, p
        self.rank = None

    def free(self):
        if self.h:
            lib().gbgpu_dense_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def rref(self, form: int = FORM_PLAIN) -> int:
        r = C.c_int32(0)
        self.ctx._check(lib().gbgpu_dense_rref(self.ctx.h, self.h, form, C.byref(r)), "dense rref")
        self.rank = int(r.value)
        return self.rank

    def result(self):
        n = self.shape[1]
        rk = self.rank or 0
        rows = np.zeros((max(rk, 1), max(n, 1)), dtype=np.uint32)
        lead = np.zeros(max(rk, 1), dtype=np.int32)
        self.ctx._check(lib().gbgpu_dense_download(self.ctx.h, self.h, rows.ctypes.data,
                                                   lead.ctypes.data), "dense download")
        return rows[:rk, :n], lead[:rk]
