"""Groebner bases over the rationals by multimodular F4 on the GPU.

The reference computes over QQ with fraction-free GMP linear algebra
(la_qq.c:271-362, export at io.c:238-330).  This module reaches the same
bases the multimodular way instead:

1. the input system, coefficients as numerator/denominator pairs
   (f4_julia's convention for field_char 0, io.c:238-330), is reduced
   modulo a sequence of 31-bit primes (largest first, skipping primes that
   divide a denominator);
2. each image runs through the F4 host with its linear algebra on the GPU
   (``Context.f4``), giving a basis with monic elements;
3. images are grouped by their support (lengths and exponents); primes whose
   support differs from the majority are unlucky and dropped;
4. the coefficients of the lucky images are combined by the Chinese
   remainder theorem and lifted to fractions by rational reconstruction
   (Wang's half-extended Euclid, bound sqrt(M / 2));
5. the lift is accepted once the image of a further prime, not used to
   form it, is reproduced exactly by it;
6. each element is written with integer coefficients: multiplied by the
   least common multiple of its denominators and divided by the content,
   so its leading coefficient is positive (the form gb's QQ output has).

The F4 runner is a callable ``run(system) -> Basis``; the product path passes
the GPU runner (:func:`gpu_runner`).  With ``run_many(systems) -> [Basis]``
(:func:`gpu_batch_runner`) the images are computed a block of primes at a
time, the F4 runs of a block in lockstep with every round's matrices of all
its images reduced in one batched launch; the images are then consumed in
prime order exactly as one at a time, so the lift and the primes it uses are
the same.  The reduced basis (reduce_gb = 1) is
unique, so its lift is independent of how each image was computed; the
non-reduced basis is the one gb's own F4 order of operations produces, which
the F4 host reproduces for every prime (its bases equal gb's over prime
fields), so its lift equals gb's QQ output up to the integer normalisation.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Callable, Iterator, List, Optional, Sequence

import numpy as np

from . import native

MR_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)  # exact below 3.3e24


def is_prime(n: int) -> bool:
    """Miller-Rabin with the first twelve primes as bases: deterministic for
    every n below 3.3e24, which covers the 31-bit primes used here."""
    if n < 2:
        return False
    for q in MR_BASES:
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in MR_BASES:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def primes_below(start: int = 2 ** 31) -> Iterator[int]:
    """Primes p < start, largest first."""
    n = start - 1
    while n > 2:
        if is_prime(n):
            yield n
        n -= 1


def rational_reconstruction(a: int, m: int) -> Optional[Fraction]:
    """The fraction n / d with n = a d (mod m), |n| <= N and 0 < d <= N for
    N = floor(sqrt(m / 2)), if one exists (it is then unique); else None.
    Wang's method: the extended Euclidean algorithm on (m, a) stopped at the
    first remainder <= N."""
    a %= m
    bound = math.isqrt(m // 2)
    r0, r1 = m, a
    s0, s1 = 0, 1
    while r1 > bound:
        q = r0 // r1
        r0, r1 = r1, r0 - q * r1
        s0, s1 = s1, s0 - q * s1
    if s1 == 0 or abs(s1) > bound or math.gcd(r1, abs(s1)) != 1:
        return None
    if s1 < 0:
        r1, s1 = -r1, -s1
    return Fraction(r1, s1)


@dataclass
class QQBasis:
    """A basis over QQ: element i has lens[i] terms, exponents exps (nterms
    x nv) and integer coefficients cfs (Python ints, each element primitive
    with a positive leading coefficient); fractions holds the monic form."""
    nv: int
    lens: List[int]
    exps: np.ndarray
    cfs: List[int]
    fractions: List[Fraction]
    primes: List[int] = field(default_factory=list)
    unlucky: List[int] = field(default_factory=list)

    @property
    def nterms(self) -> int:
        return len(self.cfs)


def _image(nv: int, lens, exps, num, den, p: int, order: int) -> native.PolySystem:
    res = [(int(n) * pow(int(d), -1, p)) % p for n, d in zip(num, den)]
    return native.PolySystem(nv=nv, p=p, lens=np.asarray(lens, np.int32),
                             exps=np.asarray(exps, np.int32).reshape(-1, nv),
                             cfs=np.asarray(res, np.int64).astype(np.int32), order=order)


def _integral(lens: Sequence[int], fr: Sequence[Fraction]) -> List[int]:
    out, t = [], 0
    for L in lens:
        el = fr[t:t + L]
        t += L
        lcm = 1
        for c in el:
            lcm = lcm * c.denominator // math.gcd(lcm, c.denominator)
        ints = [int(c * lcm) for c in el]
        g = 0
        for v in ints:
            g = math.gcd(g, v)
        g = g or 1
        if ints and ints[0] < 0:
            g = -g
        out.extend(v // g for v in ints)
    return out


def _images(run, run_many, make, primes_ok: Iterator[int]):
    """(p, image basis) in prime order: one at a time with ``run``, or in
    blocks of 3, 4, 8, 16, ... primes with ``run_many`` (two images to lift
    and one to check, then doubling)."""
    if run_many is None:
        for p in primes_ok:
            yield p, run(make(p))
        return
    size = 3
    while True:
        block = []
        for p in primes_ok:
            block.append(p)
            if len(block) == size:
                break
        if not block:
            return
        for p, b in zip(block, run_many([make(p) for p in block])):
            yield p, b
        size = 4 if size == 3 else 2 * size


def groebner_qq(run: Optional[Callable[[native.PolySystem], native.Basis]], nv: int, lens, exps,
                num, den, order: int = 0, primes: Optional[Sequence[int]] = None,
                max_primes: int = 256, run_many=None) -> QQBasis:
    """Basis over QQ of the system with coefficients num[i] / den[i] (den > 0
    or negative; zero denominators are an error), from the images the
    runner computes (``run`` one image at a time, or ``run_many`` a block of
    images at once).  Raises RuntimeError if max_primes images do not settle
    the lift."""
    num = [int(x) for x in num]
    den = [int(x) for x in den]
    if len(num) != len(den) or len(num) != int(np.sum(lens)):
        raise ValueError("one numerator and one denominator per term")
    if any(d == 0 for d in den):
        raise ValueError("zero denominator")
    source = iter(primes) if primes is not None else primes_below()

    def usable() -> Iterator[int]:
        n = 0
        for q in source:
            if n >= max_primes:
                return
            if any(d % q == 0 for d in den):
                continue
            n += 1
            yield q

    groups = {}        # support -> [modulus, CRT residues (ints), primes]
    cand = None        # (support, fractions): the current lift of the largest group
    used = 0
    for p, b in _images(run, run_many, lambda q: _image(nv, lens, exps, num, den, q, order), usable()):
        used += 1
        key = (tuple(int(x) for x in b.lens), np.asarray(b.exps, np.int32).tobytes())
        cf = [int(x) for x in b.cfs]
        if cand is not None and cand[0] == key and all(
                (c.numerator * pow(c.denominator, -1, p) - r) % p == 0 for c, r in zip(cand[1], cf)):
            # an image the lift did not come from reproduces it: accept
            g = groups[key]
            unlucky = [q for k, v in groups.items() if k != key for q in v[2]]
            return QQBasis(nv=nv, lens=list(key[0]), exps=np.asarray(b.exps, np.int32).copy(),
                           cfs=_integral(key[0], cand[1]), fractions=cand[1],
                           primes=g[2] + [p], unlucky=unlucky)
        g = groups.setdefault(key, [1, [0] * len(cf), []])
        inv = pow(g[0] % p, -1, p)
        g[1] = [x + g[0] * (((r - x) * inv) % p) for x, r in zip(g[1], cf)]
        g[0] *= p
        g[2].append(p)
        # the support seen for the most primes is taken as the lucky one
        lead = max(groups, key=lambda k: len(groups[k][2]))
        m, res, _ = groups[lead]
        fr = []
        for x in res:
            c = rational_reconstruction(x, m)
            if c is None:
                break
            fr.append(c)
        cand = (lead, fr) if len(fr) == len(res) else None
    raise RuntimeError(f"no stable rational lift after {used} primes")


def gpu_runner(ctx: "native.Context", form: int = native.FORM_MFMA, reduce_gb: int = 1):
    """F4 images on the GPU: the host of gbgpu_f4 with the LA in ``form``."""
    def run(system: native.PolySystem) -> native.Basis:
        return ctx.f4(system, form=form, reduce_gb=reduce_gb)
    return run


def gpu_batch_runner(ctx: "native.Context", form: int = native.FORM_MFMA, reduce_gb: int = 1):
    """Blocks of F4 images on the GPU in lockstep (gbgpu_f4_multimod): every
    round's matrices of the block's images in one batched launch.  The
    runner's ``stats`` list gets each block's launch counts."""
    def run_many(systems):
        bases, st = ctx.f4_multimod(systems, form=form, reduce_gb=reduce_gb)
        run_many.stats.append(dict(st, images=len(systems)))
        return bases
    run_many.stats = []
    return run_many
