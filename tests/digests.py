"""SHA-256 digests of canonical result serialisations, shared by the
generator of the full-size fixtures (tests/golden/make_digests.py, which
runs the oracle) and the GPU tests that compare against them.  Integers are
serialised little-endian with fixed widths, so equal digests mean equal
results bit for bit."""
import hashlib

import numpy as np


def _h(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype.str).encode())
        h.update(np.int64(a.size).tobytes())
        h.update(a.tobytes())
    return h.hexdigest()


def dense(rows: np.ndarray, lead: np.ndarray) -> str:
    """A dense reduced echelon form: rank x n rows in decreasing lead order
    and their lead columns."""
    return _h(np.asarray(rows, dtype="<u4"), np.asarray(lead, dtype="<i4"))


def sparse_rows(r) -> str:
    """LA result rows (gb_amd.matrix.SparseRows): row pointers, columns,
    coefficients, in the reference's output order."""
    return _h(np.asarray(r.rp, dtype="<i8"), np.asarray(r.ci, dtype="<i4"),
              np.asarray(r.cf, dtype="<u4"))


def basis(b) -> str:
    """A Groebner basis in f4_julia's export arrays (gb_amd.native.Basis)."""
    return _h(np.asarray(b.lens, dtype="<i4"), np.asarray(b.exps, dtype="<i4"),
              np.asarray(b.cfs, dtype="<u4"))
