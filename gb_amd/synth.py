"""Seeded synthetic inputs with the shapes of F4 matrices.

All generators take an explicit seed and use numpy's PCG64, so every input
of the tests and of the benchmark can be regenerated bit for bit.

* :func:`f4_shaped` -- one F4 matrix: ``npiv`` known-pivot rows with a
  staircase of leads on the left columns (unit lead, lognormal lengths),
  plus to-reduce rows whose leads sit on pivot columns, as symbolic
  preprocessing produces them (symbol.c:129, convert.c:77-79).  Rows come
  out ordered by (lead, length) like gb's sort_matrix_rows (order.c:98-131).
* :func:`dense_block` -- a uniform dense block mod p, optionally of a given
  rank (a product L x R mod p).
* :func:`batch_31bit` -- many F4-shaped matrices of uneven sizes, each mod
  its own prime just below 2^31.
"""
from __future__ import annotations

import numpy as np

from .matrix import F4Matrix, sort_rows


def _distinct(rng, lo: int, hi: int, k: int) -> np.ndarray:
    """k distinct integers from [lo, hi) in random order (k <= hi - lo)."""
    span = hi - lo
    k = min(k, span)
    if k <= 0:
        return np.zeros(0, dtype=np.int64)
    if 4 * k > span:
        return lo + rng.permutation(span)[:k]
    out = np.zeros(0, dtype=np.int64)
    while len(out) < k:
        raw = np.concatenate([out, rng.integers(lo, hi, size=k - len(out) + 8)])
        _, first = np.unique(raw, return_index=True)
        out = raw[np.sort(first)]
    return out[:k]


def f4_shaped(nr: int, nc: int, npiv: int, p: int, seed: int,
              piv_len=(50, 400), piv_mu: float = np.log(150.0),
              piv_sigma: float = 0.5, red_nnz: int | None = None,
              red_density: float | None = 0.01) -> F4Matrix:
    """One F4-shaped matrix.

    ``npiv`` pivot rows have leads 0..npiv-1 (left block = first npiv
    columns) and lognormal lengths clipped to ``piv_len``; the other
    ``nr - npiv`` rows have a lead drawn uniformly from the pivot columns and
    ``red_nnz`` entries (default ``red_density * nc``).  Non-lead columns are
    uniform to the right of the lead; coefficients uniform in [1, p-1] with a
    unit lead."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if red_nnz is None:
        red_nnz = max(2, int(round(red_density * nc)))
    nred = nr - npiv
    lens_p = np.clip(np.round(rng.lognormal(piv_mu, piv_sigma, size=npiv)),
                     piv_len[0], piv_len[1]).astype(np.int64)
    leads = np.concatenate([np.arange(npiv, dtype=np.int64),
                            rng.integers(0, npiv, size=nred)])
    lens = np.concatenate([lens_p, np.full(nred, red_nnz, dtype=np.int64)])
    lens = np.minimum(lens, nc - leads)
    rows = []
    for i in range(nr):
        rest = _distinct(rng, int(leads[i]) + 1, nc, int(lens[i]) - 1)
        rows.append(np.concatenate([[leads[i]], rest]))
    lens = np.array([len(r) for r in rows], dtype=np.int64)
    rp = np.zeros(nr + 1, dtype=np.int64)
    np.cumsum(lens, out=rp[1:])
    ci = np.concatenate(rows).astype(np.int32)
    cf = rng.integers(1, p, size=int(rp[-1]), dtype=np.int64).astype(np.uint32)
    cf[rp[:-1]] = 1
    m = F4Matrix(nr=nr, nc=nc, ncl=npiv, p=p, rp=rp, ci=ci,
                 rcf=np.arange(nr, dtype=np.int32), cp=rp.copy(), cf=cf)
    return sort_rows(m)


def _matmul_mod(a: np.ndarray, b: np.ndarray, p: int) -> np.ndarray:
    """(a @ b) mod p exactly, through float64 products whose partial sums
    stay below 2^53 (operands split into 16-bit halves when needed)."""
    k = a.shape[1]
    if k * (p - 1) ** 2 < 2 ** 53:
        return np.fmod(a.astype(np.float64) @ b.astype(np.float64), p).astype(np.uint32)
    assert k * (2 ** 16) ** 2 < 2 ** 53
    a = a.astype(np.int64)
    b = b.astype(np.int64)
    ah, al = (a >> 16).astype(np.float64), (a & 0xFFFF).astype(np.float64)
    bh, bl = (b >> 16).astype(np.float64), (b & 0xFFFF).astype(np.float64)
    hh = np.fmod(ah @ bh, p).astype(np.int64)
    mid = (np.fmod(ah @ bl, p).astype(np.int64) + np.fmod(al @ bh, p).astype(np.int64)) % p
    ll = np.fmod(al @ bl, p).astype(np.int64)
    s16 = (1 << 16) % p
    s32 = (1 << 32) % p
    return ((hh * s32 % p + mid * s16 % p + ll) % p).astype(np.uint32)


def dense_block(m: int, n: int, p: int, seed: int, rank: int | None = None) -> np.ndarray:
    """Uniform m x n block mod p; with ``rank`` it is L (m x rank) times
    R (rank x n) mod p, whose rank is ``rank`` with high probability."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if rank is None or rank >= min(m, n):
        return rng.integers(0, p, size=(m, n), dtype=np.int64).astype(np.uint32)
    left = rng.integers(0, p, size=(m, rank), dtype=np.int64)
    right = rng.integers(0, p, size=(rank, n), dtype=np.int64)
    return _matmul_mod(left, right, p)


def primes_below(bound: int, count: int) -> list[int]:
    """The ``count`` largest primes below ``bound`` (trial division by the
    primes up to sqrt(bound))."""
    lim = int(bound ** 0.5) + 1
    sieve = np.ones(lim + 1, dtype=bool)
    sieve[:2] = False
    for i in range(2, int(lim ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = False
    small = np.nonzero(sieve)[0]
    out = []
    c = bound - 1
    while len(out) < count and c > 1:
        if c % 2 and not np.any(c % small[small < c] == 0):
            out.append(int(c))
        c -= 1
    return out


def batch_31bit(count: int, seed: int, rows=(500, 20000), cols_factor: float = 1.5,
                density=(0.005, 0.02), piv_frac: float = 0.6,
                row_mu: float | None = None, row_sigma: float = 1.0):
    """``count`` F4-shaped matrices of uneven sizes, each mod its own prime
    just below 2^31: rows lognormal clipped to ``rows``, columns
    ``cols_factor`` x rows, density uniform in ``density``, ``piv_frac`` of
    the rows known pivots."""
    rng = np.random.Generator(np.random.PCG64(seed))
    primes = primes_below(2 ** 31, count)
    if row_mu is None:
        row_mu = float(np.log(np.sqrt(rows[0] * rows[1])))
    out = []
    for k in range(count):
        nr = int(np.clip(round(rng.lognormal(row_mu, row_sigma)), rows[0], rows[1]))
        nc = int(round(cols_factor * nr))
        npiv = int(round(piv_frac * nr))
        dens = float(rng.uniform(*density))
        plen = max(2, int(round(dens * nc)))
        out.append(f4_shaped(nr, nc, npiv, primes[k], seed=int(rng.integers(1 << 62)),
                             piv_len=(max(2, plen // 3), max(3, 3 * plen)),
                             piv_mu=float(np.log(plen)), piv_sigma=0.5,
                             red_density=dens))
    return out


def splitmix64(seed: int):
    """The splitmix64 generator (Steele, Lea, Flood 2014) as an iterator."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def dense_quadratic_system(n: int, p: int, seed: int, order: int = 0):
    """n quadratics in n variables with every monomial of degree <= 2
    present, coefficients uniform in [1, p - 1] from splitmix64(seed)."""
    from .native import PolySystem
    mons = [(i, j) for i in range(n) for j in range(i, n)]  # x_i x_j
    terms = []
    for i, j in mons:
        e = [0] * n
        e[i] += 1
        e[j] += 1
        terms.append(e)
    for i in range(n):
        e = [0] * n
        e[i] = 1
        terms.append(e)
    terms.append([0] * n)
    g = splitmix64(seed)
    lens, exps, cfs = [], [], []
    for _ in range(n):
        lens.append(len(terms))
        for e in terms:
            exps.extend(e)
            cfs.append(1 + next(g) % (p - 1))
    return PolySystem(nv=n, p=p, lens=np.array(lens, np.int32), exps=np.array(exps, np.int32),
                      cfs=np.array(cfs, np.int64).astype(np.int32), order=order)
