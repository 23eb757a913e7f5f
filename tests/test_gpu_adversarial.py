"""Exactness of the int8-digit matrix-core paths at the edges of their
bounds, on the GPU against the oracle, bit for bit.

A residue x enters the matrix cores as nd signed base-256 digits of its
representative in [-128 S, 127 S] (S = 1 + 256 + ... + 256^(nd-1)): x itself
when x <= 127 S, else x - p.  The extreme operands are 127 S (every digit
127), 127 S + 1 (the most negative representative, top digit -128) and
p - 1 (representative -1); the accumulators are largest when whole rows and
columns carry them.  Two shapes put such values straight into the products:

  * dense echelon form (k_update_mfma, K = 64, and the panel select): rows
    0..63 are [I | X], the rest [Y | Z], so the first panel's update is
    Z -= Y X with X, Y drawn from the extreme values;
  * the sparse pass's dense route (k_update_wide, K = 256, digits_balanced2
    for nd = 2): known pivots with unit leads and no other left entries
    (T = I) and right parts X, to-reduce rows [Y | Z]: D' = Z - Y X over 256
    pivot columns at once; a second variant adds extreme entries to the
    left parts of the pivots (non-trivial group inverses, narrow updates).

Primes cover one digit (127, 251), two (257, 32003, 65521), three (65537,
16777213) and four (2147483629, 2147483647)."""
import numpy as np
import pytest

from gb_amd import native
from gb_amd.matrix import from_dense_rows, sort_rows
from oracle import oracle

pytestmark = pytest.mark.gpu

PRIMES = [127, 251, 257, 32003, 65521, 65537, 16777213, 2147483629, 2147483647]


def _nd(p):
    return 1 if p <= 255 else 2 if p <= 65521 else 3 if p <= (1 << 24) else 4


def extremes(p):
    S = sum(256 ** i for i in range(_nd(p)))
    vals = {127 * S, 127 * S + 1, p - 1, p - 2, 1, (p - 1) // 2, (p + 1) // 2, 128, 255}
    return sorted(v for v in vals if 0 < v < p)


def _fill(rng, shape, p, mode):
    """Extreme values: cycling through the set, or one value per row."""
    ex = np.array(extremes(p), dtype=np.uint64)
    if mode == "mixed":
        return ex[rng.integers(0, len(ex), size=shape)].astype(np.uint32)
    # whole rows of one extreme value (every digit saturated for 127 S and
    # 127 S + 1)
    pick = ex[np.arange(shape[0]) % len(ex)]
    return np.repeat(pick[:, None], shape[1], axis=1).astype(np.uint32)


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("mode", ["mixed", "rows"])
@pytest.mark.parametrize("form", [native.FORM_MFMA, native.FORM_PLAIN])
def test_dense_echelon_extreme_operands(gpu_lib, p, mode, form):
    rng = np.random.default_rng(p % 1000 + (mode == "rows"))
    m, n = 256, 640
    a = np.zeros((m, n), dtype=np.uint32)
    a[:64, :64] = np.eye(64, dtype=np.uint32)
    a[:64, 64:] = _fill(rng, (64, n - 64), p, mode)
    a[64:, :] = _fill(rng, (m - 64, n), p, mode)
    print(f"dense extremes p={p} mode={mode} form={form}", flush=True)
    exp_rows, exp_lead = oracle.dense_rref(a, p, 4)
    d = gpu_lib.dense(a, p)
    assert d.rref(form) == len(exp_lead)
    rows, lead = d.result()
    assert np.array_equal(lead, exp_lead)
    assert np.array_equal(rows, exp_rows)


def _route_matrix(p, mode, tri, seed, ncl=256, ncr=600, nrl=320):
    """An F4 matrix whose sparse pass takes the dense route in the MFMA form
    (pivot rows carry most of the columns)."""
    rng = np.random.default_rng(seed)
    nc = ncl + ncr
    piv = np.zeros((ncl, nc), dtype=np.uint32)
    piv[np.arange(ncl), np.arange(ncl)] = 1
    piv[:, ncl:] = _fill(rng, (ncl, ncr), p, mode)
    if tri:
        up = np.triu(np.ones((ncl, ncl), dtype=bool), 1) & (rng.random((ncl, ncl)) < 0.3)
        vals = _fill(rng, (ncl, ncl), p, "mixed")
        piv[:, :ncl][up] = vals[up]
    red = _fill(rng, (nrl, nc), p, mode)
    leads = rng.integers(0, ncl, size=nrl)
    for i, l in enumerate(leads):
        red[i, :l] = 0
        if red[i, l] == 0:
            red[i, l] = 1
    rows = [piv[i] for i in range(ncl)] + [red[i] for i in range(nrl)]
    return sort_rows(from_dense_rows(rows, nc, ncl, p))


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("mode,tri", [("mixed", False), ("rows", False), ("mixed", True)])
def test_dense_route_extreme_operands(gpu_lib, p, mode, tri):
    m = _route_matrix(p, mode, tri, seed=p % 997 + 3 * tri)
    print(f"dense route extremes p={p} mode={mode} tri={tri}", flush=True)
    exp = oracle.la_sparse_dense(m, nthreads=4)
    for form in (native.FORM_MFMA, native.FORM_PLAIN):
        got = gpu_lib.la_ff(m, la_option=1, form=form)
        assert got.equals(exp), form


def test_sparse_pass_work_counts_match_the_oracle(gpu_lib):
    """The roofline numerator: pivot rows applied and pivot entries streamed
    are properties of the input (S(r) = supp(c_r A^-1)); the GPU's
    instrumented pass counts them as the oracle does."""
    from gb_amd import synth
    m = synth.f4_shaped(3000, 4500, 1800, 65521, seed=11, piv_len=(20, 200),
                        piv_mu=np.log(60), red_nnz=45)
    d = gpu_lib.upload(m)
    assert d.sparse_pass_work() == oracle.sparse_pass_work(m, nthreads=4)
    d.free()
