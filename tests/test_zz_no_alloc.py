"""gbgpu_dense_rref_noalloc through the C-ABI against gbgpu_dense_rref.

The allocation-free entry point runs the dense echelon form on a block the
caller holds in device memory, with all scratch in the library's static
device array.  Two calls in a row on one context, the second on a smaller
block (the array keeps the first call's contents), must give exactly the
rows and lead columns of gbgpu_dense_upload + gbgpu_dense_rref +
gbgpu_dense_download on the same block, in both forms; a block whose scratch
exceeds the array is refused before any device work.  The test is the
caller here: it allocates the device buffers itself, through the HIP
runtime libgbgpu is linked against (torch ships a runtime of its own,
whose device memory libgbgpu's runtime does not know)."""
import ctypes as C

import numpy as np
import pytest

from gb_amd import native

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime mapped for libgbgpu."""

    def __init__(self):
        native.lib()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line and "torch" not in line:
                    path = line.split()[-1]
                    break
        assert path, "libgbgpu's HIP runtime is not mapped"
        self.h = C.CDLL(path)
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(nbytes, 4)) == 0
        self.live.append(p.value)
        return p.value

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        d = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(d, arr.ctypes.data, arr.nbytes, H2D) == 0
        return d

    def get(self, d, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            assert self.h.hipMemcpy(out.ctypes.data, d, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.h.hipFree(C.c_void_p(p))
        self.live = []


def _block(m, n, p, seed, rank=None):
    rng = np.random.default_rng(seed)
    if rank is None:
        return rng.integers(0, p, size=(m, n), dtype=np.uint64).astype(np.uint32)
    # L (m x rank) R (rank x n) mod p, with R split into 16-bit halves so
    # that no uint64 sum overflows (rank < 2^10, entries < 2^31)
    L = rng.integers(0, p, size=(m, rank), dtype=np.uint64)
    R = rng.integers(0, p, size=(rank, n), dtype=np.uint64)
    lo = (L @ (R & 0xFFFF)) % p
    hi = (L @ (R >> 16)) % p
    return ((lo + hi * 65536) % p).astype(np.uint32)


def _reference(ctx, a, p, form):
    d = ctx.dense(a, p)
    rk = d.rref(form)
    rows, lead = d.result()
    d.free()
    return rk, rows.copy(), lead.copy()


def _noalloc(ctx, hip, a, p, form):
    m, n = a.shape
    mn = min(m, n)
    da = hip.put(a)
    rows = hip.put(np.full((mn, n), 0xFFFFFFFF, dtype=np.uint32))
    lead = hip.put(np.full(mn, -7, dtype=np.int32))
    rk = ctx.dense_rref_noalloc(da, m, n, p, form, rows, lead)
    return rk, hip.get(rows, (mn, n), np.uint32)[:rk], hip.get(lead, mn, np.int32)[:rk]


CASES = [
    # (first block, second smaller block, p, form)
    (dict(m=700, n=900, seed=1), dict(m=300, n=500, seed=2), 65521, native.FORM_MFMA),
    (dict(m=640, n=800, seed=3), dict(m=200, n=260, seed=4), 2147483629, native.FORM_PLAIN),
    (dict(m=500, n=700, seed=5, rank=300), dict(m=128, n=200, seed=6), 251, native.FORM_MFMA),
    (dict(m=400, n=600, seed=7), dict(m=65, n=70, seed=8, rank=40), 2147483629, native.FORM_MFMA),
]


@pytest.mark.parametrize("first,second,p,form", CASES)
def test_noalloc_matches_allocating_entry(first, second, p, form):
    ctx = native.Context(0)
    hip = _Hip()
    try:
        for spec in (first, second):
            a = _block(spec["m"], spec["n"], p, spec["seed"], spec.get("rank"))
            print(f"no-alloc dense rref: {a.shape} p={p} form={form}", flush=True)
            rk, rows, lead = _reference(ctx, a, p, form)
            if spec.get("rank") is not None:
                assert rk == spec["rank"]
            rk2, rows2, lead2 = _noalloc(ctx, hip, a, p, form)
            assert rk2 == rk
            assert np.array_equal(lead2, lead)
            assert np.array_equal(rows2, rows)
    finally:
        hip.free()
        ctx.close()


def test_noalloc_refuses_blocks_beyond_the_scratch():
    L = native.lib()
    m = n = 100000
    need = L.gbgpu_dense_rref_noalloc_bytes(m, n, 65521, native.FORM_MFMA)
    assert need > 64 << 20
    ctx = native.Context(0)
    hip = _Hip()
    try:
        buf = hip.alloc(64)
        with pytest.raises(native.GpuError, match=r"\(-5\)"):
            ctx.dense_rref_noalloc(buf, m, n, 65521, native.FORM_MFMA, buf, buf)
    finally:
        hip.free()
        ctx.close()
