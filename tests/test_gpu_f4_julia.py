"""gbgpu_f4_julia through the C-ABI, called the way GB.jl calls gb's f4_julia
(f4.c:190-212): the same argument list, the basis returned in malloc'ed
arrays (bld, blen, bexp, bcf) that the caller frees.  The bases must equal
the ones gb's own tests assert (test/cyclic_5_drl_17_bit.c,
test/red_gb_17_bit.c, test/cyclic_6_drl_17_bit_la42_t2.c,
test/cyclic_4_lex_17_bit.c), element for element; la_option 101 selects the
plain form, every other code the MFMA form."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from gb_amd import native

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "f4_golden.json")))
CASES = ["cyclic_5_drl_17_bit", "red_gb_17_bit", "cyclic_6_drl_17_bit_la42_t2", "cyclic_4_lex_17_bit"]


def _f4_julia(c, la_option):
    L = native.lib()
    fn = L.gbgpu_f4_julia
    i32, pi32 = C.c_int32, C.POINTER(C.c_int32)
    fn.restype = C.c_int64
    fn.argtypes = [pi32, C.POINTER(pi32), C.POINTER(pi32), C.POINTER(C.c_void_p), pi32, pi32,
                   C.c_void_p] + [i32] * 12
    lens = np.ascontiguousarray(c["lens"], np.int32)
    exps = np.ascontiguousarray(c["exps"], np.int32)
    cfs = np.ascontiguousarray(c["cfs"], np.int32)
    bld = C.c_int32(0)
    blen, bexp, bcf = pi32(), pi32(), C.c_void_p()
    nterms = fn(C.byref(bld), C.byref(blen), C.byref(bexp), C.byref(bcf),
                lens.ctypes.data_as(pi32), exps.ctypes.data_as(pi32), C.c_void_p(cfs.ctypes.data),
                c["field_char"], c["mon_order"], c["nr_vars"], c["nr_gens"], 17, c["nr_threads"],
                0, 0, la_option, c["reduce_gb"], 0, 0)
    assert nterms > 0
    n = bld.value
    got_len = np.ctypeslib.as_array(blen, shape=(n,)).copy()
    got_exp = np.ctypeslib.as_array(bexp, shape=(nterms * c["nr_vars"],)).copy()
    got_cf = np.ctypeslib.as_array(C.cast(bcf, pi32), shape=(nterms,)).copy()
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    for p in (C.cast(blen, C.c_void_p), C.cast(bexp, C.c_void_p), bcf):
        libc.free(p)
    return int(nterms), got_len, got_exp, got_cf


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("la_option", [101, 102])
def test_f4_julia_drop_in(gpu_lib, name, la_option):
    c = GOLDEN[name]
    print(f"gbgpu_f4_julia {name} la_option {la_option}", flush=True)
    nterms, lens, exps, cfs = _f4_julia(c, la_option)
    assert nterms == len(c["tcf"])
    assert list(lens) == c["tlen"]
    assert list(exps) == c["texp"]
    assert list(cfs) == c["tcf"]
