"""GPU parity at the benchmark configurations' full sizes, against digests of
the oracle's results (tests/golden/full_size_digests.json, written by
tests/golden/make_digests.py in a CPU container; the oracle needs minutes
there, so the GPU tests compare SHA-256 digests of canonical serialisations,
which are equal exactly when the results are equal bit for bit).

  * the 20000 x 30000 F4 matrix mod 65521: la_option 1 in both forms, and
    the sparse pass's work counters (the roofline numerator) equal the
    oracle's count of the same implementation-independent quantities;
  * the 4096 x 8192 dense blocks mod 251, 65521, 2147483629 and the
    rank-3000 block: the reduced echelon form in both forms (64 panels,
    multi-tile selects, prunes and the look-ahead at these sizes);
  * the 64-matrix batch, rows 500-20000, each modulo its own 31-bit prime:
    every matrix's rows in both forms, first call and graph replay;
  * F4 end to end on 14 (and 12) dense quadratics in 14 variables mod
    65521: the basis and, round by round, the matrix shape, the new rows and
    a hash of the round's echelon form, in both forms."""
import json
import os

import numpy as np
import pytest

from gb_amd import native, synth
from tests import digests

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "full_size_digests.json")))
FORMS = [native.FORM_PLAIN, native.FORM_MFMA]


def test_c2_la_option1_and_work_counts(gpu_lib):
    c = FIX["c2"]
    m = synth.f4_shaped(c["nr"], c["nc"], c["npiv"], c["p"], seed=c["seed"])
    d = gpu_lib.upload(m)
    nd, napp, nent = d.sparse_pass_work()
    assert (nd, napp, nent) == (c["nonzero_remainders"], c["pivots_applied"], c["pivot_entries"])
    for form in FORMS:
        print(f"c2 la_option 1, form {form}", flush=True)
        assert d.la_sparse_dense(form) == c["new_pivots"]
        assert digests.sparse_rows(d.rows()) == c["la_option1"]
    d.free()


C3_KEYS = sorted(k for k in FIX["c3"] if k != "shape")


@pytest.mark.parametrize("key", C3_KEYS)
@pytest.mark.parametrize("form", FORMS)
def test_c3_dense_block(gpu_lib, key, form):
    c = FIX["c3"][key]
    sh = FIX["c3"]["shape"]
    rank = 3000 if key.endswith("_rank3000") else None
    a = synth.dense_block(sh["m"], sh["n"], c["p"], seed=sh["seed"], rank=rank)
    print(f"c3 {key} form {form}", flush=True)
    d = gpu_lib.dense(a, c["p"])
    assert d.rref(form) == c["rank"]
    rows, lead = d.result()
    assert digests.dense(rows, lead) == c["rref"]
    d.free()


@pytest.fixture(scope="module")
def c4_mats():
    c = FIX["c4"]
    mats = synth.batch_31bit(c["count"], seed=c["seed"])
    assert [int(m.p) for m in mats] == [x["p"] for x in c["matrices"]]
    assert [int(m.nr) for m in mats] == [x["nr"] for x in c["matrices"]]
    return mats


@pytest.mark.parametrize("form", FORMS)
def test_c4_batch_64_matrices(gpu_lib, c4_mats, form):
    c = FIX["c4"]
    b = gpu_lib.batch(c4_mats)
    for call in range(2):  # layout, tuning and capture; then a replay
        print(f"c4 batch form {form} call {call}", flush=True)
        npv = b.la(form)
        for i, x in enumerate(c["matrices"]):
            assert npv[i] == x["new_pivots"], i
            assert digests.sparse_rows(b.rows(i)) == x["la_option1"], i
    b.free()


@pytest.mark.parametrize("nv", [12, 14])
@pytest.mark.parametrize("form", [native.FORM_MFMA, native.FORM_PLAIN])
def test_c5_f4_end_to_end(gpu_lib, nv, form):
    c = FIX["c5"][f"vars{nv}"]
    s = synth.dense_quadratic_system(nv, 65521, seed=1)
    print(f"c5 F4 {nv} variables form {form}", flush=True)
    b = gpu_lib.f4(s, form=form)
    got = [list(r.values()) for r in b.stats["trace"]]
    assert got == c["trace"]
    assert len(b.lens) == c["elements"] and int(np.sum(b.lens)) == c["terms"]
    assert digests.basis(b) == c["basis"]
