"""Groebner bases over QQ with every image's F4 linear algebra on the GPU
(gb_amd/qq.py with the GPU runner): gb's QQ golden bases, and the same lift
as the all-CPU run."""
import json
import math
import os
from fractions import Fraction

import pytest

from gb_amd import qq
from gb_amd.native import FORM_MFMA, FORM_PLAIN
from oracle import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "qq_golden.json")))


def _solve(case, run):
    n = case["nr_gens"]
    return qq.groebner_qq(run, case["nr_vars"], case["lens"][:n], case["exps"], case["num"],
                          case["den"], order=case["mon_order"])


@pytest.mark.parametrize("form", [FORM_PLAIN, FORM_MFMA])
def test_gpu_qq_example_equals_gb(gpu_lib, form):
    case = GOLD["qq_example"]
    b = _solve(case, qq.gpu_runner(gpu_lib, form, reduce_gb=case["reduce_gb"]))
    assert len(b.lens) == case["bld"] and b.nterms == case["nterms"]
    assert b.cfs == case["tcf"]


@pytest.mark.parametrize("form", [FORM_PLAIN, FORM_MFMA])
def test_gpu_red_gb_qq_equals_gb_up_to_element_scaling(gpu_lib, form):
    case = GOLD["red_gb_qq"]
    b = _solve(case, qq.gpu_runner(gpu_lib, form, reduce_gb=1))
    assert len(b.lens) == case["bld"] and b.nterms == case["nterms"]
    t = 0
    for L in b.lens:
        e, g = b.cfs[t:t + L], case["tcf"][t:t + L]
        r = Fraction(g[0], e[0])
        assert [Fraction(x) * r for x in e] == [Fraction(x) for x in g]
        assert e[0] > 0 and math.gcd(*e) == 1
        t += L
    # the same lift as the CPU images give
    c = _solve(case, lambda s: oracle.f4(s, nthreads=4, reduce_gb=1))
    assert c.cfs == b.cfs and c.lens == b.lens and c.primes == b.primes


@pytest.mark.parametrize("form", [FORM_PLAIN, FORM_MFMA])
def test_gpu_qq_images_through_the_batch(gpu_lib, form):
    """The images of a block of primes run in lockstep, every round's
    matrices of all of them in one batched launch (gbgpu_f4_multimod): the
    same lift from the same primes as one image at a time, and gb's basis."""
    for name in ("qq_example", "red_gb_qq"):
        case = GOLD[name]
        rg = case["reduce_gb"] if name == "qq_example" else 1
        run_many = qq.gpu_batch_runner(gpu_lib, form, reduce_gb=rg)
        n = case["nr_gens"]
        b = qq.groebner_qq(None, case["nr_vars"], case["lens"][:n], case["exps"], case["num"],
                           case["den"], order=case["mon_order"], run_many=run_many)
        one = _solve(case, qq.gpu_runner(gpu_lib, form, reduce_gb=rg))
        assert b.cfs == one.cfs and b.lens == one.lens and b.primes == one.primes
        if name == "qq_example":
            assert b.cfs == case["tcf"]
        st = run_many.stats[0]
        assert st["images"] == 3 and st["batches"] > 0
        assert st["matrices"] == 3 * st["batches"]  # the lucky images share every round
