"""Brute-force modular linear algebra in Python integers, independent of the
oracle's code path: used to pin the oracle on small inputs."""
from __future__ import annotations


def rref(rows, p):
    """Gauss-Jordan over GF(p) of a list of integer rows.  Returns the
    nonzero rows of the reduced row echelon form, ascending lead."""
    m = [[x % p for x in r] for r in rows]
    if not m:
        return []
    n = len(m[0])
    piv_rows = []
    r = 0
    for c in range(n):
        sel = None
        for i in range(r, len(m)):
            if m[i][c]:
                sel = i
                break
        if sel is None:
            continue
        m[r], m[sel] = m[sel], m[r]
        inv = pow(m[r][c], p - 2, p)
        m[r] = [(x * inv) % p for x in m[r]]
        for i in range(len(m)):
            if i != r and m[i][c]:
                f = m[i][c]
                m[i] = [(a - f * b) % p for a, b in zip(m[i], m[r])]
        piv_rows.append(r)
        r += 1
        if r == len(m):
            break
    return [m[i] for i in range(r)]


def lead(row):
    for j, x in enumerate(row):
        if x:
            return j
    return -1


def inverse_unit_upper(a, p):
    """Inverse of a unit upper-triangular matrix, one column at a time by
    back substitution (a different route from row elimination)."""
    n = len(a)
    inv = [[0] * n for _ in range(n)]
    for col in range(n):
        e = [1 if i == col else 0 for i in range(n)]
        x = [0] * n
        for i in range(n - 1, -1, -1):
            s = e[i]
            for k in range(i + 1, n):
                s -= a[i][k] * x[k]
            x[i] = s % p
        for i in range(n):
            inv[i][col] = x[i]
    return inv
