"""Pin the CPU oracle (oracle/liboracle.so) before trusting it.

Checks, in order: the reference's own known-answer test, closed forms of
the modular inverse, the field-size thresholds of the LA selector, and the
canonical-RREF property of every LA routine against brute-force
Gauss-Jordan in Python integers (tests/brute.py) on small F4-shaped inputs,
plus an explicit-inverse route (x = c_left * A^-1) for the sparse pass.
"""
import json
import os
import random

import numpy as np
import pytest

from gb_amd import synth
from gb_amd.matrix import from_dense_rows, sort_rows
from oracle import oracle
from tests import brute

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PRIMES = [101, 127, 251, 32003, 65521, 131071, 2147483629, 2147483647]


def test_normalize_row_known_answer():
    kat = json.load(open(os.path.join(GOLD, "normalize_row.json")))
    row = [kat["row0"]] + [300 - i for i in range(1, kat["len"])]
    out = oracle.normalize_dense_row(row, kat["fc"])
    assert out[0] == 1
    for i, v in kat["expect"].items():
        assert out[int(i)] == v


@pytest.mark.parametrize("p", PRIMES)
def test_mod_inverse_closed_form(p):
    rng = random.Random(p)
    assert oracle.mod_inverse(0, p) == 0
    for a in [1, 2, p - 1, p - 2] + [rng.randrange(1, p) for _ in range(200)]:
        inv = oracle.mod_inverse(a, p)
        assert 0 <= inv < p
        assert (a * inv) % p == 1
        assert inv == pow(a, p - 2, p)  # Fermat


def test_field_size_thresholds():
    # io.c:815-829 (8 / 16 / 32 storage) and io.c:710-729 (17 vs 31 bit)
    cases = {2: 8, 101: 8, 127: 8, 128: 16, 251: 16, 32003: 16, 32767: 16,
             32768: 17, 65521: 17, 131071: 17, 131072: 31, 2147483647: 31}
    for p, bits in cases.items():
        assert oracle.arith_bits(p) == bits


def _small_f4(seed, p, nr=48, nc=70, npiv=26, red=12):
    return synth.f4_shaped(nr, nc, npiv, p, seed=seed, piv_len=(2, 12),
                           piv_mu=np.log(6), piv_sigma=0.6, red_nnz=red)


def _expected_output(m):
    """Right-hand RREF rows in decreasing lead order (brute force)."""
    dense = [list(map(int, r)) for r in m.to_dense()]
    rows = [r for r in brute.rref(dense, m.p) if brute.lead(r) >= m.ncl]
    return rows[::-1]


def _as_lists(sr, nc):
    return [list(map(int, r)) for r in sr.to_dense(nc)]


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_la_options_match_brute_force(p, seed):
    m = _small_f4(seed, p)
    expect = _expected_output(m)
    for fn in (oracle.la_sparse_dense, oracle.la_sparse):
        for th in (1, 4):
            got = fn(m, nthreads=th)
            assert _as_lists(got, m.nc) == expect
            lens = np.diff(got.rp)
            assert np.all(got.cf[got.rp[:-1]] == 1)
            for i in range(got.np):
                c = got.ci[got.rp[i]:got.rp[i + 1]]
                assert np.all(np.diff(c) > 0) and c[0] >= m.ncl
            assert np.all(np.diff(got.leads()) < 0)  # decreasing lead
            assert np.all(lens > 0)


@pytest.mark.parametrize("p", [251, 65521, 2147483629])
def test_rank_deficient_and_zero_rows(p):
    # to-reduce rows that are combinations of pivot rows reduce to zero
    rng = np.random.default_rng(7)
    nc, ncl = 30, 12
    piv = []
    for c in range(ncl):
        r = np.zeros(nc, np.int64)
        r[c] = 1
        nz = rng.choice(np.arange(c + 1, nc), size=4, replace=False)
        r[nz] = rng.integers(1, p, size=4)
        piv.append(r)
    red = []
    for k in range(10):
        coef = rng.integers(0, p, size=ncl)
        coef[: k % 5] = 0
        coef[k % 5] = 1
        r = np.zeros(nc, np.int64)
        for j in range(ncl):
            r = (r + int(coef[j]) * piv[j]) % p
        if k % 3 == 0:  # add one right-hand direction
            r[ncl + k] = (r[ncl + k] + 1) % p
        red.append(r)
    m = sort_rows(from_dense_rows(piv + red, nc, ncl, p))
    expect = _expected_output(m)
    assert len(expect) == 4
    for fn in (oracle.la_sparse_dense, oracle.la_sparse):
        assert _as_lists(fn(m), nc) == expect


@pytest.mark.parametrize("p", [127, 32003, 65521, 2147483647])
def test_sparse_pass_explicit_inverse(p):
    """D' rows equal c_right - (c_left A^-1) B, with A^-1 computed by back
    substitution instead of row-by-row elimination."""
    m = _small_f4(11, p, nr=40, nc=50, npiv=20, red=9)
    dense = m.to_dense().astype(object)
    leads = m.leads()
    first = {}
    upiv = []
    for i, l in enumerate(leads):
        if int(l) not in first:
            first[int(l)] = i
        else:
            upiv.append(i)
    upiv = upiv[::-1]
    ncl = m.ncl
    A = [[int(dense[first[r]][c]) for c in range(ncl)] for r in range(ncl)]
    B = [[int(dense[first[r]][c]) for c in range(ncl, m.nc)] for r in range(ncl)]
    Ainv = brute.inverse_unit_upper(A, p)
    expect = []
    for i in upiv:
        cl = [int(x) for x in dense[i][:ncl]]
        x = [sum(cl[k] * Ainv[k][j] for k in range(ncl)) % p for j in range(ncl)]
        d = [(int(dense[i][ncl + j]) - sum(x[k] * B[k][j] for k in range(ncl))) % p
             for j in range(m.ncr)]
        if any(d):
            expect.append((i, d))
    got, src = oracle.sparse_pass(m)
    assert [int(s) for s in src] == [i for i, _ in expect]
    assert [list(map(int, r)) for r in got] == [d for _, d in expect]


@pytest.mark.parametrize("p", [101, 251, 65521, 2147483629])
@pytest.mark.parametrize("shape,rank", [((30, 45), None), ((40, 25), None), ((36, 50), 11)])
def test_dense_rref_brute_force(p, shape, rank):
    a = synth.dense_block(shape[0], shape[1], p, seed=5, rank=rank)
    expect = brute.rref([list(map(int, r)) for r in a], p)[::-1]
    for th in (1, 3):
        got, lead = oracle.dense_rref(a, p, nthreads=th)
        assert [list(map(int, r)) for r in got] == expect
        assert list(lead) == [brute.lead(r) for r in expect]
    if rank is not None:
        assert len(expect) == rank


@pytest.mark.parametrize("p", [101, 65521, 2147483647])
def test_interreduce_brute_force(p):
    rng = np.random.default_rng(3)
    nc = 40
    leads = sorted(rng.choice(nc - 5, size=12, replace=False))
    rows = []
    for l in leads:
        r = np.zeros(nc, np.int64)
        r[l] = 1
        nz = rng.choice(np.arange(l + 1, nc), size=min(6, nc - l - 1), replace=False)
        r[nz] = rng.integers(1, p, size=len(nz))
        rows.append(r)
    m = from_dense_rows(rows, nc, 0, p)
    got = oracle.interreduce(m)
    expect = brute.rref([list(map(int, r)) for r in rows], p)
    assert _as_lists(got, nc) == expect
    assert np.all(np.diff(got.leads()) > 0)


def test_contract_violation_is_an_error():
    # a lead outside the known-pivot block is rejected, not undefined
    m = from_dense_rows([[1, 2, 0, 0], [0, 0, 1, 3]], 4, 1, 101)
    with pytest.raises(ValueError):
        oracle.la_sparse_dense(m)
