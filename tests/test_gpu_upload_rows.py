"""gbgpu_mat_upload_rows: a matrix handed over as monomial ids plus a column
map gives, bit for bit, the la_option 1 result of the same matrix handed
over in gb's CSR (gbgpu_mat_upload), and the oracle's; ids outside the map
and rows outside the id array are refused."""
import numpy as np
import pytest

from gb_amd import synth
from gb_amd.native import FORM_MFMA, FORM_PLAIN, GpuError
from oracle import oracle

pytestmark = pytest.mark.gpu


def _by_ids(m, seed):
    """m's rows as monomial ids: columns get distinct ids in a map three
    times wider than the matrix (the rest of the map holds junk that no
    row refers to), rows sit in mon in a shuffled order with junk between
    them."""
    rng = np.random.default_rng(seed)
    nmap = 3 * m.nc + 7
    ids = rng.choice(nmap, size=m.nc, replace=False).astype(np.int32)
    colmap = rng.integers(-(1 << 30), 1 << 30, size=nmap).astype(np.int32)
    colmap[ids] = np.arange(m.nc, dtype=np.int32)
    order = rng.permutation(m.nr)
    src = np.zeros(m.nr, np.int64)
    parts, pos = [], 0
    for q in order:
        gap = int(rng.integers(0, 4))
        parts.append(rng.integers(0, nmap, size=gap).astype(np.int32))
        pos += gap
        src[q] = pos
        row = m.ci[m.rp[q]:m.rp[q + 1]]
        parts.append(ids[row])
        pos += len(row)
    mon = np.concatenate(parts) if parts else np.zeros(0, np.int32)
    return mon, src, colmap


@pytest.mark.parametrize("p", [251, 65521, 2147483647])
@pytest.mark.parametrize("form", [FORM_PLAIN, FORM_MFMA])
def test_upload_rows_same_result_as_csr(gpu_lib, p, form):
    m = synth.f4_shaped(1500, 2600, 900, p, seed=11, piv_len=(5, 120), piv_mu=np.log(40),
                        red_nnz=90)
    mon, src, colmap = _by_ids(m, seed=3)
    d = gpu_lib.upload_rows(m, mon, src, colmap)
    n = d.la_sparse_dense(form)
    got = d.rows()
    d.free()
    ref = gpu_lib.la_ff(m, la_option=1, form=form)
    exp = oracle.la_sparse_dense(m, nthreads=4)
    assert n == exp.np
    assert got.equals(ref)
    assert got.equals(exp)


def test_upload_rows_large_rows(gpu_lib):
    """rows of thousands of entries (several passes of a workgroup per row)"""
    m = synth.f4_shaped(3000, 40000, 1800, 65521, seed=4, piv_len=(200, 6000),
                        piv_mu=np.log(2000), red_nnz=4000)
    mon, src, colmap = _by_ids(m, seed=9)
    d = gpu_lib.upload_rows(m, mon, src, colmap)
    d.la_sparse_dense(FORM_MFMA)
    got = d.rows()
    d.free()
    assert got.equals(oracle.la_sparse_dense(m, nthreads=8))


def test_upload_rows_refuses_bad_ids(gpu_lib):
    m = synth.f4_shaped(200, 300, 120, 65521, seed=2, red_nnz=20)
    mon, src, colmap = _by_ids(m, seed=1)
    bad = mon.copy()
    bad[src[5]] = len(colmap)  # one id past the map
    with pytest.raises(GpuError, match="column index out of range"):
        gpu_lib.upload_rows(m, bad, src, colmap)
    short = mon[:int(src.max())]  # the last row runs past the id array
    with pytest.raises(GpuError, match="outside the monomial-id array"):
        gpu_lib.upload_rows(m, short, src, colmap)
    # the context stays usable
    d = gpu_lib.upload_rows(m, mon, src, colmap)
    d.la_sparse_dense(FORM_PLAIN)
    assert d.rows().equals(oracle.la_sparse_dense(m))
    d.free()


@pytest.mark.parametrize("p", [65521, 2147483647])
def test_upload_rows_with_coefficient_store(gpu_lib, p):
    """coefficient arrays taken from a device store (appended in pieces,
    junk between them): the same result as the host pool"""
    m = synth.f4_shaped(1200, 2000, 700, p, seed=8, piv_len=(5, 100), piv_mu=np.log(30),
                        red_nnz=60)
    mon, src, colmap = _by_ids(m, seed=5)
    rng = np.random.default_rng(2)
    store = gpu_lib.cfstore(p)
    ncf = len(m.cp) - 1
    cp = np.zeros(ncf + 1, np.int64)
    store.append(rng.integers(0, p, size=5))
    for i in range(ncf):
        cp[i] = store.append(m.cf[m.cp[i]:m.cp[i + 1]])
        store.append(rng.integers(0, p, size=int(rng.integers(0, 3))))
    cp[ncf] = store.size()
    for form in (FORM_PLAIN, FORM_MFMA):
        d = gpu_lib.upload_rows(m, mon, src, colmap, store=store, cp=cp)
        d.la_sparse_dense(form)
        got = d.rows()
        d.free()
        assert got.equals(gpu_lib.la_ff(m, la_option=1, form=form))
    # refused: an array past the store's end, an entry >= p, another prime
    bad = cp.copy()
    bad[-1] = store.size() + 1
    with pytest.raises(GpuError, match="outside the store"):
        gpu_lib.upload_rows(m, mon, src, colmap, store=store, cp=bad)
    with pytest.raises(GpuError, match="not reduced"):
        store.append(np.array([1, p], np.uint32))
    other = gpu_lib.cfstore(251)
    with pytest.raises(GpuError, match="another context or prime"):
        gpu_lib.upload_rows(m, mon, src, colmap, store=other, cp=cp)
    other.free()
    store.free()
