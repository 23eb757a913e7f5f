"""The F4 host with every round's linear algebra on the GPU: gb's golden
bases bit for bit, and the same bases as the all-CPU run (oracle linear
algebra) on dense quadratic systems."""
import json
import os

import numpy as np
import pytest

from gb_amd import native, synth
from gb_amd.native import FORM_MFMA, FORM_PLAIN
from oracle import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "f4_golden.json")))
CASES = sorted(k for k, v in GOLDEN.items() if "tlen" in v)
FORMS = [FORM_PLAIN, FORM_MFMA]


def system_of(c):
    return native.PolySystem(nv=c["nr_vars"], p=c["field_char"], lens=np.array(c["lens"]),
                             exps=np.array(c["exps"]), cfs=np.array(c["cfs"]),
                             order=c["mon_order"])


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("name", CASES)
def test_gpu_f4_matches_gb_golden_basis(gpu_lib, name, form):
    c = GOLDEN[name]
    b = gpu_lib.f4(system_of(c), form=form, reduce_gb=c["reduce_gb"])
    assert list(b.lens) == c["tlen"]
    assert list(b.cfs) == c["tcf"]
    assert list(b.exps.reshape(-1)) == c["texp"]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("n,p", [(8, 32003), (9, 2147483647), (10, 65521), (8, 251)])
def test_gpu_f4_dense_quadratic_matches_cpu_run(gpu_lib, n, p, form):
    s = synth.dense_quadratic_system(n, p, seed=3)
    exp = oracle.f4(s, nthreads=8)
    got = gpu_lib.f4(s, form=form)
    assert np.array_equal(got.lens, exp.lens)
    assert np.array_equal(got.exps, exp.exps)
    assert np.array_equal(got.cfs, exp.cfs)
    assert got.stats["rounds"] == exp.stats["rounds"]
