"""The gb drop-in adapters (include/gbgpu_gb.h) driven like gb's F4 loop
drives its linear_algebra / interreduce_matrix_rows pointers, checked
against the oracle and against gb's ownership and statistics contract."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from gb_amd import synth
from gb_amd.matrix import SparseRows, from_dense_rows
from oracle import oracle

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(__file__), "_bin", "adapter_driver")


def _run(m, mode, width, laopt):
    nrl = m.nr - len(np.unique(m.leads()))
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([mode, width, m.nr, m.nc, m.ncl, len(m.cp) - 1, m.p, laopt, nrl],
                     np.int32).tofile(f)
            m.rp.astype(np.int64).tofile(f)
            m.ci.astype(np.int32).tofile(f)
            m.rcf.astype(np.int32).tofile(f)
            m.cp.astype(np.int64).tofile(f)
            m.cf.astype(np.uint32).tofile(f)
        r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = open(fout, "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    rp = np.frombuffer(raw[4:4 + 8 * (n + 1)], np.int64).copy()
    nnz = int(rp[-1])
    off = 4 + 8 * (n + 1)
    ci = np.frombuffer(raw[off:off + 4 * nnz], np.int32).copy()
    cf = np.frombuffer(raw[off + 4 * nnz:off + 8 * nnz], np.uint32).copy()
    return SparseRows(rp=rp, ci=ci, cf=cf)


@pytest.mark.parametrize("p,width", [(101, 8), (32003, 16), (65521, 32), (2147483647, 32)])
@pytest.mark.parametrize("laopt", [101, 1])
def test_linear_algebra_adapter(p, width, laopt):
    m = synth.f4_shaped(600, 900, 350, p, seed=width + p % 7, piv_len=(4, 40),
                        piv_mu=np.log(12), red_nnz=25)
    assert _run(m, 0, width, laopt).equals(oracle.la_sparse_dense(m))


@pytest.mark.parametrize("p,width", [(101, 8), (32003, 16), (2147483647, 32)])
def test_interreduce_adapter(p, width):
    rng = np.random.default_rng(p % 89)
    nc = 200
    leads = np.sort(rng.choice(nc - 5, size=50, replace=False))
    rows = []
    for l in leads:
        r = np.zeros(nc, np.int64)
        r[l] = 1
        nz = rng.choice(np.arange(l + 1, nc), size=min(8, nc - l - 1), replace=False)
        r[nz] = rng.integers(1, p, size=len(nz))
        rows.append(r)
    m = from_dense_rows(rows, nc, 0, p)
    assert _run(m, 1, width, 1).equals(oracle.interreduce(m))
