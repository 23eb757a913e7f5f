"""The F4 host (gbgpu_f4) end to end on the CPU, with the oracle's linear
algebra in every round: the bases must equal, element for element and
coefficient for coefficient, the bases that gb's own tests assert
(test/cyclic_*.c, test/red_gb_*.c; fixtures in tests/golden)."""
import json
import os

import numpy as np
import pytest

from gb_amd import native, synth
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "f4_golden.json")))
CASES = sorted(k for k, v in GOLDEN.items() if "tlen" in v)


def system_of(c):
    return native.PolySystem(nv=c["nr_vars"], p=c["field_char"], lens=np.array(c["lens"]),
                             exps=np.array(c["exps"]), cfs=np.array(c["cfs"]),
                             order=c["mon_order"])


def assert_golden(b, c):
    assert list(b.lens) == c["tlen"]
    assert list(b.cfs) == c["tcf"]
    assert list(b.exps.reshape(-1)) == c["texp"]


@pytest.mark.parametrize("name", CASES)
def test_f4_host_matches_gb_golden_basis(name):
    c = GOLDEN[name]
    b = oracle.f4(system_of(c), nthreads=2, reduce_gb=c["reduce_gb"])
    assert_golden(b, c)


def test_f4_host_reduced_basis_is_reduced_form_of_golden():
    """The reduced basis is unique: reducing gb's (non-reduced) cyclic-6
    basis gives the same polynomials as running F4 with reduce_gb."""
    c = GOLDEN["cyclic_6_drl_17_bit_la1_t1"]
    mine = oracle.f4(system_of(c), nthreads=2, reduce_gb=1)
    nv = c["nr_vars"]
    gold = native.PolySystem(nv=nv, p=c["field_char"], lens=np.array(c["tlen"]),
                             exps=np.array(c["texp"]), cfs=np.array(c["tcf"]), order=0)
    again = oracle.f4(gold, nthreads=2, reduce_gb=1)
    assert list(mine.lens) == list(again.lens)
    assert np.array_equal(mine.exps, again.exps)
    assert np.array_equal(mine.cfs, again.cfs)
    # monic, terms strictly decreasing (DRL), no lead divides another lead
    t = 0
    leads = []
    for L in mine.lens:
        assert mine.cfs[t] == 1
        leads.append(mine.exps[t])
        t += L
    for i, a in enumerate(leads):
        for j, bb in enumerate(leads):
            assert i == j or not np.all(a <= bb)


def test_dense_quadratic_system_is_seeded():
    a = synth.dense_quadratic_system(8, 32003, seed=1)
    b = synth.dense_quadratic_system(8, 32003, seed=1)
    assert np.array_equal(a.cfs, b.cfs) and len(a.lens) == 8 and a.lens[0] == 45
    assert a.cfs.min() >= 1 and a.cfs.max() <= 32002


def test_f4_host_dense_quadratic_8_vars():
    """8 quadratics in 8 variables (every monomial of degree <= 2): a
    zero-dimensional ideal with 2^8 solutions over the algebraic closure
    (generic Bezout number), so the reduced DRL basis leaves exactly 256
    standard monomials."""
    s = synth.dense_quadratic_system(8, 32003, seed=1)
    b = oracle.f4(s, nthreads=4, reduce_gb=1)
    leads = []
    t = 0
    for L in b.lens:
        leads.append(tuple(b.exps[t]))
        t += L
    # count standard monomials (not divisible by any lead) up to degree 9
    from itertools import combinations_with_replacement
    L = np.array(leads)
    nstd = 0
    for d in range(10):
        for c in combinations_with_replacement(range(8), d):
            e = np.bincount(np.array(c, dtype=np.int64), minlength=8)
            if not np.any(np.all(L <= e, axis=1)):
                nstd += 1
    assert nstd == 256


def test_f4_host_exponents_beyond_one_byte():
    """{x^100 y^150 + x^240, x^200 y^150 + 1} mod 65521, DRL: the S-pair
    multiplies x^240 by x^100, so exponents pass 255 during the run.  Since
    x^100 (x^100 y^150 + x^240) - (x^200 y^150 + 1) = x^340 - 1, x is a
    unit and y^150 + x^140 = x^-100 f1 lies in the ideal; the two leads
    x^340 and y^150 are coprime, so the reduced basis is exactly
    {x^340 - 1, y^150 + x^140}."""
    p = 65521
    s = native.PolySystem(nv=2, p=p, lens=np.array([2, 2]),
                          exps=np.array([100, 150, 240, 0, 200, 150, 0, 0]),
                          cfs=np.array([1, 1, 1, 1]), order=0)
    b = oracle.f4(s, nthreads=2, reduce_gb=1)
    got = set()
    t = 0
    for L in b.lens:
        got.add(tuple((tuple(int(v) for v in b.exps[t + a]), int(b.cfs[t + a])) for a in range(L)))
        t += L
    assert got == {((( 340, 0), 1), ((0, 0), p - 1)), (((0, 150), 1), ((140, 0), 1))}


def test_f4_host_exponent_overflow_is_an_error():
    """Lex with y > x: y - x^30000 and x^5000 y + 1.  The first lead
    divides the second, and the pair's row x^5000 (y - x^30000) holds
    x^35000, beyond gb's int16 exponent range (data.h:63): the host
    reports GBGPU_ERANGE (-5) instead of wrapping the exponent."""
    s = native.PolySystem(nv=2, p=65521, lens=np.array([2, 2]),
                          exps=np.array([1, 0, 0, 30000, 1, 5000, 0, 0]),
                          cfs=np.array([1, -1, 1, 1]), order=1)
    with pytest.raises(native.GpuError, match=r"\(-5\)"):
        oracle.f4(s, nthreads=2)
    # an input exponent past the range is refused the same way
    s = native.PolySystem(nv=1, p=65521, lens=np.array([1]), exps=np.array([40000]),
                          cfs=np.array([1]), order=0)
    with pytest.raises(native.GpuError, match=r"\(-5\)"):
        oracle.f4(s, nthreads=1)


# katsura-12 mod 65521, DRL (test/katsura_12_drl_17_bit.c): per round the
# degree, pairs selected, matrix rows, columns, pivot columns, nonzeros and
# new basis elements.  gb's own info_level 2 trace of this system (the
# reference test compiled from source) gives for the degree-9 round a
# 27707 x 27798 matrix at 1.204 % density, 2064 rows to reduce and 120 new
# elements; rows 27707 - 25643 = 2064, 9273235 / (27707 * 27798) = 1.2040 %.
KATSURA_12_TRACE = [
    (2, 2, 24, 91, 22, 282, 2),
    (3, 14, 202, 433, 188, 3095, 14),
    (4, 85, 878, 1404, 793, 39031, 51),
    (5, 407, 3599, 4339, 3192, 323015, 124),
    (6, 1106, 9972, 10562, 8866, 1509066, 211),
    (7, 1983, 19363, 19447, 17380, 4282762, 252),
    (8, 2436, 26010, 25764, 23574, 7544434, 210),
    (9, 2064, 27707, 27798, 25643, 9273235, 120),
    (10, 1191, 25796, 26696, 24605, 8341827, 45),
    (11, 449, 24580, 26188, 24131, 7234038, 10),
    (12, 100, 22508, 24456, 22408, 6064426, 1),
    (13, 10, 21378, 23415, 21368, 4865985, 0),
]


def test_f4_host_katsura_12_matrices_per_round():
    """The host builds gb's matrices round by round on katsura-12: same
    pairs (first pair of equal lcm kept, update.c:164-183 and
    hash.c:351-389), same input order (order.c:63-95), same reducers
    (symbol.c:151-200).  The degree-9 row is gb's measured shape."""
    c = GOLDEN["katsura_12_drl_17_bit"]
    b = oracle.f4(system_of(c), nthreads=8, reduce_gb=0)
    got = [tuple(r.values())[:7] for r in b.stats["trace"]]
    assert got == KATSURA_12_TRACE
    r9 = got[7]
    assert (r9[2], r9[3], r9[2] - r9[4], r9[6]) == (27707, 27798, 2064, 120)
    assert round(100.0 * r9[5] / (r9[2] * r9[3]), 3) == 1.204
    assert len(b.lens) == 1050 and int(np.sum(b.lens)) == 1481052


@pytest.mark.parametrize("name", ["cyclic_6_drl_17_bit_la1_t1", "cyclic_5_lex_17_bit",
                                  "red_gb_17_bit"])
def test_f4_host_more_than_64_variables(name):
    """gb's golden systems placed into a ring of 70 variables, the k-th
    variable sent to an increasing index set that mixes indices below 64
    with indices 64 and up (whose divisibility-mask bits v & 63 coincide
    with the low ones'), the others unused: the monomial order on the used
    variables is unchanged, so the basis is the golden basis with its
    variables sent the same way."""
    c = GOLDEN[name]
    nv = c["nr_vars"]
    big = 70
    where = np.array(([0, 1, 2, 64, 65, 66, 67])[:nv])
    assert np.all(np.diff(where) > 0)

    def spread(flat, nterms):
        e = np.zeros((nterms, big), np.int64)
        e[:, where] = np.array(flat).reshape(nterms, nv)
        return e.reshape(-1)

    n = sum(c["lens"])
    s = native.PolySystem(nv=big, p=c["field_char"], lens=np.array(c["lens"]),
                          exps=spread(c["exps"], n), cfs=np.array(c["cfs"]),
                          order=c["mon_order"])
    b = oracle.f4(s, nthreads=2, reduce_gb=c["reduce_gb"])
    assert list(b.lens) == c["tlen"]
    assert list(b.cfs) == c["tcf"]
    assert np.array_equal(b.exps.reshape(-1), spread(c["texp"], sum(c["tlen"])))
