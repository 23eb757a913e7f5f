"""Host-side F4 matrix container: the flat CSR wire format of the C-ABI.

gb keeps an F4 matrix as ``mat_t`` (src/data.h:137-155): an array of row
pointers ``hm_t *r[i]`` whose header is ``[cf-array id, len % 4, len]``
followed by ``len`` column indices (lead first, rest unordered,
src/convert.c:111-119), and coefficient arrays ``bs->cf_32[id]`` shared by
all rows that are multiples of one basis element (src/hash.c:903-926).

The flat form below carries exactly that information without the row
pointers: ``rp``/``ci`` is a CSR of column indices, ``rcf[i]`` is the
coefficient-array id of row ``i`` and ``cp``/``cf`` hold the coefficient
arrays back to back.  ``include/gbgpu.h`` declares the same layout as
``gbgpu_csr_t``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class F4Matrix:
    """One F4 matrix: ``nr`` rows, ``nc`` columns, the first ``ncl`` of
    which are the known-pivot (left, "A|C") columns; p is the field size."""

    nr: int
    nc: int
    ncl: int
    p: int
    rp: np.ndarray   # int64[nr + 1]
    ci: np.ndarray   # int32[nnz]
    rcf: np.ndarray  # int32[nr]
    cp: np.ndarray   # int64[ncf + 1]
    cf: np.ndarray   # uint32[cp[-1]]

    @property
    def ncr(self) -> int:
        return self.nc - self.ncl

    @property
    def nnz(self) -> int:
        return int(self.rp[-1])

    def leads(self) -> np.ndarray:
        return self.ci[self.rp[:-1]]

    def row(self, i: int):
        """(columns, coefficients) of row i."""
        b, e = int(self.rp[i]), int(self.rp[i + 1])
        c0 = int(self.cp[self.rcf[i]])
        return self.ci[b:e], self.cf[c0:c0 + (e - b)]

    def check(self) -> None:
        """Validate the wire-format invariants the LA routines rely on."""
        assert self.rp.dtype == np.int64 and self.ci.dtype == np.int32
        assert self.rcf.dtype == np.int32 and self.cp.dtype == np.int64
        assert self.cf.dtype == np.uint32
        assert len(self.rp) == self.nr + 1 and self.rp[0] == 0
        assert np.all(np.diff(self.rp) > 0), "empty rows are not allowed"
        assert len(self.rcf) == self.nr
        lens = np.diff(self.rp)
        alen = np.diff(self.cp)
        assert np.all(alen[self.rcf] >= lens), "coefficient array too short"
        assert self.ci.min(initial=0) >= 0 and self.ci.max(initial=-1) < self.nc
        assert np.all(self.cf < self.p)

    def to_dense(self) -> np.ndarray:
        """Dense (nr x nc) uint32 copy, for small test matrices."""
        out = np.zeros((self.nr, self.nc), dtype=np.uint32)
        for i in range(self.nr):
            c, v = self.row(i)
            out[i, c] = v
        return out


@dataclass
class SparseRows:
    """Rows returned by the LA routines: CSR with strictly ascending columns
    and unit lead coefficients (the hm_t rows + mat->cf_32 arrays that gb's
    linear_algebra hands back, la_ff_32.c:1666-1736)."""

    rp: np.ndarray   # int64[np + 1]
    ci: np.ndarray   # int32[nnz]
    cf: np.ndarray   # uint32[nnz]

    @property
    def np(self) -> int:
        return len(self.rp) - 1

    def leads(self) -> np.ndarray:
        return self.ci[self.rp[:-1]] if self.np else np.zeros(0, np.int32)

    def equals(self, other: "SparseRows") -> bool:
        return (np.array_equal(self.rp, other.rp)
                and np.array_equal(self.ci, other.ci)
                and np.array_equal(self.cf, other.cf))

    def to_dense(self, nc: int) -> np.ndarray:
        out = np.zeros((self.np, nc), dtype=np.uint32)
        for i in range(self.np):
            b, e = int(self.rp[i]), int(self.rp[i + 1])
            out[i, self.ci[b:e]] = self.cf[b:e]
        return out


def from_dense_rows(rows, nc: int, ncl: int, p: int) -> F4Matrix:
    """Build an F4Matrix from a list of dense rows (each its own coefficient
    array).  Column order within a row is ascending, which satisfies the
    lead-first contract."""
    rp = [0]
    ci = []
    cf = []
    for r in rows:
        nz = np.nonzero(np.asarray(r) % p)[0]
        ci.extend(nz.tolist())
        cf.extend((np.asarray(r)[nz] % p).tolist())
        rp.append(len(ci))
    rp = np.asarray(rp, dtype=np.int64)
    return F4Matrix(nr=len(rows), nc=nc, ncl=ncl, p=p, rp=rp,
                    ci=np.asarray(ci, dtype=np.int32),
                    rcf=np.arange(len(rows), dtype=np.int32),
                    cp=rp.copy(), cf=np.asarray(cf, dtype=np.uint32))


def sort_rows(m: F4Matrix) -> F4Matrix:
    """Order rows by (lead column, length), as gb's sort_matrix_rows does
    before linear algebra (order.c:98-131, f4.c:300)."""
    lens = np.diff(m.rp)
    order = np.lexsort((lens, m.leads()))
    return permute_rows(m, order)


def permute_rows(m: F4Matrix, order: np.ndarray) -> F4Matrix:
    """Rows order[0], order[1], ... of m (a permutation or a subset)."""
    order = np.asarray(order, dtype=np.int64)
    lens = np.diff(m.rp)[order]
    rp = np.zeros(len(order) + 1, dtype=np.int64)
    np.cumsum(lens, out=rp[1:])
    starts = m.rp[:-1][order]
    idx = np.repeat(starts - rp[:-1], lens) + np.arange(rp[-1], dtype=np.int64)
    return F4Matrix(nr=len(order), nc=m.nc, ncl=m.ncl, p=m.p, rp=rp,
                    ci=m.ci[idx].astype(np.int32), rcf=m.rcf[order].astype(np.int32),
                    cp=m.cp, cf=m.cf)
