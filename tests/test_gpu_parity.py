"""GPU parity: every LA entry point of libgbgpu against the CPU oracle, bit
for bit, on seeded inputs; full-size shapes through size-independent
properties (sampled rows, Freivalds projections)."""
import numpy as np
import pytest

from gb_amd import synth
from gb_amd.matrix import F4Matrix, from_dense_rows, permute_rows, sort_rows
from gb_amd.native import FORM_MFMA, FORM_PLAIN, GpuError
from oracle import oracle

pytestmark = pytest.mark.gpu

PRIMES = [127, 251, 32003, 65521, 131071, 2147483629, 2147483647]
FORMS = [FORM_PLAIN, FORM_MFMA]


def _f4(nr, nc, npiv, p, seed, **kw):
    return synth.f4_shaped(nr, nc, npiv, p, seed=seed, **kw)


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("form", FORMS)
def test_sparse_pass_small(gpu_lib, p, form):
    m = _f4(300, 500, 180, p, seed=p % 1000, piv_len=(3, 40), piv_mu=np.log(12), red_nnz=25)
    d = gpu_lib.upload(m)
    nd = d.sparse_pass(form)
    got, src = d.sparse_pass_result(nd)
    exp, esrc = oracle.sparse_pass(m)
    assert nd == len(exp)
    assert np.array_equal(src, esrc)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("p", [251, 65521, 2147483647])
@pytest.mark.parametrize("form", FORMS)
def test_sparse_pass_many_windows(gpu_lib, p, form):
    # 24000 columns: 12 column windows of the LDS accumulator
    m = _f4(900, 24000, 600, p, seed=5, piv_len=(20, 200), piv_mu=np.log(80), red_nnz=120)
    d = gpu_lib.upload(m)
    nd = d.sparse_pass(form)
    got, src = d.sparse_pass_result(nd)
    exp, esrc = oracle.sparse_pass(m, nthreads=8)
    assert np.array_equal(src, esrc)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("p", [251, 65521, 131071, 16777213, 2147483647])
@pytest.mark.parametrize("shape", [(300, 420, 200), (1500, 2600, 900), (620, 900, 600),
                                   (1000, 1200, 700)])
def test_sparse_pass_dense_route(gpu_lib, p, shape):
    """Pivot rows carrying ~10 % of the columns: the MFMA form reduces on
    the matrix cores (blocked elimination on a dense copy); same D' rows,
    same order, as the oracle, and the same la_option 1 result.  Shapes:
    one partial panel; several panels; 20 rows to reduce; a last panel of
    three groups, the third partial."""
    nr, nc, npiv = shape
    m = _f4(nr, nc, npiv, p, seed=7, piv_len=(nc // 30, nc // 5), piv_mu=np.log(nc / 10),
            red_nnz=nc // 10)
    d = gpu_lib.upload(m)
    nd = d.sparse_pass(FORM_MFMA)
    got, src = d.sparse_pass_result(nd)
    exp, esrc = oracle.sparse_pass(m, nthreads=8)
    assert np.array_equal(src, esrc)
    assert np.array_equal(got, exp)
    exp1 = oracle.la_sparse_dense(m, nthreads=8)
    assert d.la_sparse_dense(FORM_MFMA) == exp1.np
    assert d.rows().equals(exp1)


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("form", FORMS)
def test_la_option1_medium(gpu_lib, p, form):
    m = _f4(2000, 3000, 1200, p, seed=17, piv_len=(10, 120), piv_mu=np.log(40), red_nnz=40)
    exp = oracle.la_sparse_dense(m, nthreads=8)
    d = gpu_lib.upload(m)
    n = d.la_sparse_dense(form)
    got = d.rows()
    assert n == exp.np
    assert got.equals(exp)
    # la_option 2 has the same canonical output
    assert d.la_sparse(form) == exp.np
    assert d.rows().equals(exp)


@pytest.mark.parametrize("la_option", [1, 2, 42, 43, 44, 0, 7])
def test_la_ff_one_call_all_options(gpu_lib, la_option):
    m = _f4(400, 700, 250, 65521, seed=3, piv_len=(5, 50), piv_mu=np.log(15), red_nnz=20)
    exp = oracle.la_sparse(m)
    got = gpu_lib.la_ff(m, la_option=la_option, form=FORM_MFMA)
    assert got.equals(exp)


@pytest.mark.parametrize("p", [251, 65521, 131071, 16777213, 2147483629])
@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("shape,rank", [((64, 100), None), ((200, 330), None), ((330, 200), None),
                                        ((256, 520), 100), ((130, 70), 69)])
def test_dense_rref(gpu_lib, p, form, shape, rank):
    a = synth.dense_block(shape[0], shape[1], p, seed=shape[0] + shape[1], rank=rank)
    exp, elead = oracle.dense_rref(a, p, nthreads=8)
    d = gpu_lib.dense(a, p)
    r = d.rref(form)
    got, lead = d.result()
    assert r == len(exp)
    assert np.array_equal(lead, elead)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("p", [251, 65521, 2147483647])
@pytest.mark.parametrize("form", FORMS)
def test_dense_rref_structured(gpu_lib, p, form):
    """Rows zero on the first panels, dead (all-zero) rows and dependent
    rows: exercises candidate tiles that find no pivot and the pruning of
    rows that are zero for good."""
    rng = np.random.default_rng(4)
    m, n = 700, 900
    a = rng.integers(0, p, size=(m, n), dtype=np.int64)
    a[90:, :300] = 0
    a[300:420, :] = 0
    a[600:680] = (a[0:80] * 3 + a[80:160] * 5) % p
    a = a.astype(np.uint32)
    exp, elead = oracle.dense_rref(a, p, nthreads=8)
    d = gpu_lib.dense(a, p)
    r = d.rref(form)
    got, lead = d.result()
    assert r == len(exp)
    assert np.array_equal(lead, elead)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("p", [251, 65521])
def test_dense_rref_tall_strip_update(gpu_lib, p):
    """6400 rows (4096 or more): the MFMA form's side-stream update for one
    and two digits walks 256-column strips (k_update_wide) instead of 32 x 64
    tiles."""
    a = synth.dense_block(6400, 800, p, seed=21)
    exp, elead = oracle.dense_rref(a, p, nthreads=8)
    d = gpu_lib.dense(a, p)
    r = d.rref(FORM_MFMA)
    got, lead = d.result()
    d.free()
    assert r == len(exp)
    assert np.array_equal(lead, elead)
    assert np.array_equal(got, exp)


def test_dense_rref_many_contexts():
    """Six contexts alive at once (twelve streams, more than the process's
    hardware queues, so some streams share one): the MFMA panel loop of each,
    ordered by device words whose waits are all for work queued before the
    waiting kernel, finishes and matches the oracle."""
    from gb_amd import native
    p = 65521
    a = synth.dense_block(300, 700, p, seed=11)
    exp, elead = oracle.dense_rref(a, p, nthreads=8)
    ctxs = [native.Context(0) for _ in range(6)]
    try:
        for k, c in enumerate(ctxs):
            d = c.dense(a, p)
            r = d.rref(FORM_MFMA)
            got, lead = d.result()
            d.free()
            assert r == len(exp), k
            assert np.array_equal(lead, elead), k
            assert np.array_equal(got, exp), k
    finally:
        for c in ctxs:
            c.close()


@pytest.mark.parametrize("p", [251, 65521, 16777213, 2147483629])
@pytest.mark.parametrize("shape,rank", [((330, 200), None), ((256, 520), 100), ((700, 900), None)])
def test_dense_rref_event_order(gpu_lib, monkeypatch, p, shape, rank):
    """The MFMA form's panel loop with its two streams ordered by events
    (GBGPU_DENSE_SYNC=events, read at every call) instead of device words:
    the same echelon form as the oracle's, and calls in the default order
    before and after it still match."""
    a = synth.dense_block(shape[0], shape[1], p, seed=shape[0] + 7, rank=rank)
    exp, elead = oracle.dense_rref(a, p, nthreads=8)
    for mode in ("words", "events", "words"):
        if mode == "events":
            monkeypatch.setenv("GBGPU_DENSE_SYNC", "events")
        else:
            monkeypatch.delenv("GBGPU_DENSE_SYNC", raising=False)
        d = gpu_lib.dense(a, p)
        r = d.rref(FORM_MFMA)
        got, lead = d.result()
        d.free()
        assert r == len(exp), mode
        assert np.array_equal(lead, elead), mode
        assert np.array_equal(got, exp), mode


@pytest.mark.parametrize("p", [101, 65521, 2147483647])
def test_interreduce(gpu_lib, p):
    rng = np.random.default_rng(p % 97)
    nc = 300
    leads = np.sort(rng.choice(nc - 10, size=80, replace=False))
    rows = []
    for l in leads:
        r = np.zeros(nc, np.int64)
        r[l] = 1
        nz = rng.choice(np.arange(l + 1, nc), size=min(12, nc - l - 1), replace=False)
        r[nz] = rng.integers(1, p, size=len(nz))
        rows.append(r)
    m = from_dense_rows(rows, nc, 0, p)
    exp = oracle.interreduce(m)
    for form in FORMS:
        d = gpu_lib.upload(m)
        assert d.interreduce(form) == exp.np
        assert d.rows().equals(exp)


def test_edge_no_rows_to_reduce(gpu_lib):
    # every row is a known pivot: nothing to reduce, no new pivots
    m = _f4(50, 90, 50, 65521, seed=1, piv_len=(2, 10), piv_mu=np.log(5))
    d = gpu_lib.upload(m)
    assert d.sparse_pass(FORM_PLAIN) == 0
    assert d.sparse_pass(FORM_MFMA) == 0
    assert d.la_sparse_dense(FORM_MFMA) == 0
    assert d.rows().np == 0


def test_edge_everything_reduces_to_zero(gpu_lib):
    p = 32003
    rng = np.random.default_rng(9)
    nc, ncl = 60, 25
    piv = []
    for c in range(ncl):
        r = np.zeros(nc, np.int64)
        r[c] = 1
        nz = rng.choice(np.arange(c + 1, nc), size=5, replace=False)
        r[nz] = rng.integers(1, p, size=5)
        piv.append(r)
    red = []
    for k in range(30):
        coef = rng.integers(0, p, size=ncl)
        coef[k % ncl] = 1
        coef[: k % ncl] = 0
        red.append(sum(int(coef[j]) * piv[j] for j in range(ncl)) % p)
    m = sort_rows(from_dense_rows(piv + red, nc, ncl, p))
    d = gpu_lib.upload(m)
    assert d.sparse_pass(FORM_PLAIN) == 0
    assert d.sparse_pass(FORM_MFMA) == 0
    assert d.la_sparse_dense(FORM_PLAIN) == 0
    assert oracle.la_sparse_dense(m).np == 0


def test_contract_violation_is_an_error(gpu_lib):
    m = from_dense_rows([[1, 2, 0, 0], [0, 0, 1, 3]], 4, 1, 101)
    d = gpu_lib.upload(m)
    with pytest.raises(GpuError):
        d.la_sparse_dense(FORM_PLAIN)
    bad = F4Matrix(nr=1, nc=4, ncl=1, p=101, rp=np.array([0, 2], np.int64),
                   ci=np.array([0, 9], np.int32), rcf=np.zeros(1, np.int32),
                   cp=np.array([0, 2], np.int64), cf=np.array([1, 5], np.uint32))
    with pytest.raises(GpuError):
        gpu_lib.upload(bad)  # column out of range: rejected by the upload, before any compute kernel


def test_shared_coefficient_arrays(gpu_lib):
    # rows that are multiples of one basis element share one coefficient
    # array (hash.c:903-926): same result as with private copies
    m = _f4(400, 600, 240, 65521, seed=21, piv_len=(4, 30), piv_mu=np.log(10), red_nnz=18)
    lens = np.diff(m.rp)
    # make rows with equal length share the first such row's array
    first_of_len = {}
    rcf = m.rcf.copy()
    cf = m.cf.copy()
    for i in range(m.nr):
        L = int(lens[i])
        if L in first_of_len:
            rcf[i] = rcf[first_of_len[L]]
        else:
            first_of_len[L] = i
    m2 = F4Matrix(nr=m.nr, nc=m.nc, ncl=m.ncl, p=m.p, rp=m.rp, ci=m.ci, rcf=rcf.astype(np.int32),
                  cp=m.cp, cf=cf)
    exp = oracle.la_sparse_dense(m2)
    got = gpu_lib.la_ff(m2, la_option=1, form=FORM_PLAIN)
    assert got.equals(exp)


def _pivot_rows(m):
    seen = set()
    piv = []
    for i, l in enumerate(m.leads()):
        if int(l) not in seen:
            seen.add(int(l))
            piv.append(i)
    return np.array(piv)


@pytest.mark.parametrize("p", [65521, 2147483629])
@pytest.mark.parametrize("form", FORMS)
def test_full_size_sparse_pass_sampled(gpu_lib, p, form):
    """20000 x 30000 F4 matrix (12000 known pivots): the GPU pass over all
    8000 to-reduce rows, checked against the oracle on 48 sampled rows."""
    m = _f4(20000, 30000, 12000, p, seed=2024)
    d = gpu_lib.upload(m)
    nd = d.sparse_pass(form)
    got, src = d.sparse_pass_result(nd)
    piv = _pivot_rows(m)
    nonpiv = np.setdiff1d(np.arange(m.nr), piv)
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(nonpiv, size=48, replace=False))
    keep = np.sort(np.concatenate([piv, sample]))
    sub = permute_rows(m, keep)
    exp, esrc = oracle.sparse_pass(sub, nthreads=8)
    by_src = {int(s): i for i, s in enumerate(src)}
    for k, s in enumerate(esrc):
        full = int(keep[s])
        assert full in by_src, "oracle row nonzero but GPU row zero"
        assert np.array_equal(got[by_src[full]], exp[k])
    zero_sampled = set(int(x) for x in sample) - set(int(keep[s]) for s in esrc)
    assert not (zero_sampled & set(by_src)), "GPU row nonzero but oracle row zero"


@pytest.mark.parametrize("p", [65521, 2147483647])
@pytest.mark.parametrize("form", FORMS)
def test_wide_matrix_more_than_63_windows(gpu_lib, p, form):
    """600,000 columns: more than 63 windows of the widest (8192-column)
    window the sparse pass uses, so the pivot layout is bucketed into 74
    windows (gb itself has no column limit)."""
    m = _f4(400, 600000, 250, p, seed=5, red_nnz=300)
    exp = oracle.la_sparse_dense(m, nthreads=4)
    got = gpu_lib.la_ff(m, la_option=1, form=form)
    assert got.equals(exp)
