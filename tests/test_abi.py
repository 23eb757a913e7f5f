"""CPU-side checks of the C-ABI library: it loads, exports every entry point
include/gbgpu.h declares, and its host-only logic behaves (no GPU needed)."""
import ctypes as C
import os

import pytest

from gb_amd import native


def test_header_declares_the_boundary():
    names = native.declared_symbols()
    for must in ("gbgpu_init", "gbgpu_mat_upload", "gbgpu_sparse_pass", "gbgpu_la_sparse_dense",
                 "gbgpu_la_sparse", "gbgpu_interreduce", "gbgpu_la_ff", "gbgpu_dense_rref"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = native.lib()
    missing = [n for n in native.declared_symbols() if not hasattr(L, n)]
    assert missing == []


def test_arith_bits_match_reference_selector():
    L = native.lib()
    # io.c:815-829 and io.c:710-729
    for p, bits in {101: 8, 127: 8, 128: 16, 251: 16, 32003: 16, 32768: 17,
                    65521: 17, 131071: 17, 131072: 31, 2147483647: 31}.items():
        assert L.gbgpu_arith_bits(p) == bits


def test_no_device_is_an_error_not_a_fallback():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and _has_gpu():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert native.lib().gbgpu_init(0, C.byref(h)) != 0
    with pytest.raises(native.GpuError):
        native.Context(0)


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False
