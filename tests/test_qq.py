"""Groebner bases over QQ by multimodular F4 (gb_amd/qq.py), on the CPU:
the images come from the F4 host with the oracle's linear algebra.

Pinned by gb's own QQ tests (tests/golden/qq_golden.json, extracted from
test/qq_example.c and test/red_gb_qq.c): the non-reduced basis must equal
gb's output coefficient for coefficient; the reduced basis must equal gb's
element for element up to a rational factor (gb's reduced QQ elements keep
the scalings of its fraction-free interreduction, la_qq.c), with our
elements primitive and of positive leading coefficient.  Rational
reconstruction and the prime source are pinned by known answers and brute
force."""
import json
import math
import os
import random
from fractions import Fraction
from itertools import islice

import pytest

from gb_amd import qq
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "qq_golden.json")))


def _cpu_runner(reduce_gb):
    return lambda s: oracle.f4(s, nthreads=4, reduce_gb=reduce_gb)


def _solve(case, run, primes=None):
    n = case["nr_gens"]
    return qq.groebner_qq(run, case["nr_vars"], case["lens"][:n], case["exps"], case["num"],
                          case["den"], order=case["mon_order"], primes=primes)


def _elements(lens, cfs):
    out, t = [], 0
    for L in lens:
        out.append(cfs[t:t + L])
        t += L
    return out


def test_is_prime_against_trial_division():
    def slow(n):
        return n >= 2 and all(n % d for d in range(2, math.isqrt(n) + 1))
    for n in range(0, 20000):
        assert qq.is_prime(n) == slow(n)
    # Carmichael numbers and strong pseudoprimes to small bases
    for n in (561, 1105, 1729, 2047, 3215031751, 2152302898747):
        assert not qq.is_prime(n)


def test_primes_below_two_to_31():
    # the largest primes below 2^31 (2^31 - 1 is the Mersenne prime)
    assert list(islice(qq.primes_below(), 4)) == [2147483647, 2147483629, 2147483587, 2147483579]


def test_rational_reconstruction_known_answers():
    rng = random.Random(5)
    ps = list(islice(qq.primes_below(), 3))
    m = ps[0] * ps[1] * ps[2]
    bound = math.isqrt(m // 2)
    for _ in range(2000):
        n = rng.randint(-bound, bound)
        d = rng.randint(1, bound)
        f = Fraction(n, d)
        a = f.numerator * pow(f.denominator, -1, m) % m
        assert qq.rational_reconstruction(a, m) == f
    # integers and their negatives
    assert qq.rational_reconstruction(5, 101) == Fraction(5)
    assert qq.rational_reconstruction(96, 101) == Fraction(-5)
    # 1/2 mod 101 = 51
    assert qq.rational_reconstruction(51, 101) == Fraction(1, 2)
    # brute force over every residue of small moduli (primes and a product):
    # the fraction with |n|, d <= sqrt(m / 2) congruent to a, or None
    for m in (101, 1009, 2003, 31 * 37):
        nb = math.isqrt(m // 2)
        for a in range(m):
            cands = {Fraction(n, d) for n in range(-nb, nb + 1) for d in range(1, nb + 1)
                     if math.gcd(n, d) == 1 and math.gcd(d, m) == 1 and (n - a * d) % m == 0}
            assert len(cands) <= 1
            got = qq.rational_reconstruction(a, m)
            assert got == (cands.pop() if cands else None), (m, a)


def test_qq_example_equals_gb():
    """test/qq_example.c: 5 generators in 5 variables over QQ, DRL, not
    reduced: 23 elements, 325 terms, every coefficient as gb asserts."""
    case = GOLD["qq_example"]
    b = _solve(case, _cpu_runner(case["reduce_gb"]))
    assert len(b.lens) == case["bld"] and b.nterms == case["nterms"]
    assert b.cfs == case["tcf"]
    assert not b.unlucky


def test_red_gb_qq_equals_gb_up_to_element_scaling():
    """test/red_gb_qq.c: the reduced basis, 23 elements and 318 terms; each
    element proportional to gb's, ours primitive with positive lead."""
    case = GOLD["red_gb_qq"]
    b = _solve(case, _cpu_runner(case["reduce_gb"]))
    assert len(b.lens) == case["bld"] and b.nterms == case["nterms"]
    ours = _elements(b.lens, b.cfs)
    gold = _elements(b.lens, case["tcf"])
    scaled = 0
    for e, g in zip(ours, gold):
        r = Fraction(g[0], e[0])
        assert [Fraction(x) * r for x in e] == [Fraction(x) for x in g]
        assert e[0] > 0 and math.gcd(*e) == 1
        scaled += r != 1
    assert scaled == 4  # elements 6 and 12 negated, 7 times 4, 17 times 27 in gb's output
    # the monic forms are the lifted fractions
    t = 0
    for e in ours:
        assert b.fractions[t] == 1
        assert [Fraction(x, e[0]) for x in e] == b.fractions[t:t + len(e)]
        t += len(e)


def test_unlucky_prime_is_dropped():
    """A coefficient divisible by the first prime changes that image's
    support: the prime is set aside and the lift is the one the other
    primes give."""
    case = dict(GOLD["qq_example"])
    num = list(case["num"])
    num[0] = 2147483647 * 3  # first term of the first generator vanishes mod 2^31 - 1
    case["num"] = num
    run = _cpu_runner(case["reduce_gb"])
    b = _solve(case, run)
    c = _solve(case, run, primes=islice(qq.primes_below(2147483647), 64))
    assert b.unlucky == [2147483647]
    assert b.cfs == c.cfs and b.lens == c.lens


def test_denominators_and_errors():
    # x - 1/3, y + 2/5 in two variables, DRL: the reduced basis is itself
    lens = [2, 2]
    exps = [1, 0, 0, 0, 0, 1, 0, 0]
    b = qq.groebner_qq(_cpu_runner(1), 2, lens, exps, [1, -1, 1, 2], [1, 3, 1, 5])
    assert b.lens == [2, 2]
    assert sorted(map(tuple, _elements(b.lens, b.fractions))) == sorted(
        [(Fraction(1), Fraction(-1, 3)), (Fraction(1), Fraction(2, 5))])
    assert sorted(map(tuple, _elements(b.lens, b.cfs))) == sorted([(3, -1), (5, 2)])
    with pytest.raises(ValueError):
        qq.groebner_qq(_cpu_runner(1), 2, lens, exps, [1, -1, 1, 2], [1, 0, 1, 5])
    with pytest.raises(ValueError):
        qq.groebner_qq(_cpu_runner(1), 2, lens, exps, [1, -1, 1], [1, 3, 1])


def test_blocks_of_images_give_the_same_lift():
    """Images computed a block of primes at a time (the batched runner's
    shape: blocks of 3, 4, 8, ...) are consumed in prime order exactly as
    one at a time: the same lift from the same primes, and the first block
    already settles it (two images to lift, one to check)."""
    case = GOLD["qq_example"]
    one = _solve(case, _cpu_runner(case["reduce_gb"]))
    blocks = []

    def run_many(systems):
        blocks.append(len(systems))
        return [oracle.f4(s, nthreads=4, reduce_gb=case["reduce_gb"]) for s in systems]

    n = case["nr_gens"]
    many = qq.groebner_qq(None, case["nr_vars"], case["lens"][:n], case["exps"], case["num"],
                          case["den"], order=case["mon_order"], run_many=run_many)
    assert many.cfs == one.cfs and many.lens == one.lens and many.primes == one.primes
    assert blocks == [3]
    # another sequence of primes lifts to the same basis
    many = qq.groebner_qq(None, case["nr_vars"], case["lens"][:n], case["exps"], case["num"],
                          case["den"], order=case["mon_order"], run_many=run_many,
                          primes=list(islice(qq.primes_below(2 ** 30), 10)))
    assert many.cfs == one.cfs
