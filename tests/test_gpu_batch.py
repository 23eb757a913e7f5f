"""GPU parity of the batched multimodular launch: every matrix of a batch,
each modulo its own prime, against the oracle's la_option 1 on that matrix
alone, bit for bit; first call (layout, tuning, graph capture) and graph
replays alike."""
import numpy as np
import pytest

from gb_amd import synth
from gb_amd.matrix import from_dense_rows, sort_rows
from gb_amd.native import FORM_MFMA, FORM_PLAIN, GpuError
from oracle import oracle

pytestmark = pytest.mark.gpu

FORMS = [FORM_PLAIN, FORM_MFMA]


def _small_batch(primes, seed):
    mats = []
    for k, p in enumerate(primes):
        nr = 150 + 70 * k
        mats.append(synth.f4_shaped(nr, int(nr * 1.5), int(nr * 0.6), p, seed=seed + k,
                                    piv_len=(3, 30), piv_mu=np.log(10), red_nnz=12))
    return mats


@pytest.mark.parametrize("form", FORMS)
def test_batch_mixed_primes(gpu_lib, form):
    primes = [2147483647, 2147483629, 65521, 251, 32003, 2147483587]
    mats = _small_batch(primes, 3)
    exp = [oracle.la_sparse_dense(m) for m in mats]
    b = gpu_lib.batch(mats)
    for _ in range(3):  # the first call captures the graph, the others replay it
        npv = b.la(form)
        for i in range(len(mats)):
            assert npv[i] == exp[i].np
            assert b.rows(i).equals(exp[i])
    info = b.info()
    assert info["replays"] == 3 and info["matrices"] == len(mats)
    assert 10 <= info["window_shift"] <= 13


@pytest.mark.parametrize("primes", [[65521, 32003, 257, 40009, 251], [251, 241, 239, 2]],
                         ids=["two_digits", "one_digit"])
@pytest.mark.parametrize("form", FORMS)
def test_batch_small_primes(gpu_lib, primes, form):
    """The batch's digit count follows its largest prime: two digits (the
    single-precision reduction of the products) and one digit."""
    mats = _small_batch(primes, 9)
    exp = [oracle.la_sparse_dense(m) for m in mats]
    b = gpu_lib.batch(mats)
    for _ in range(2):
        npv = b.la(form)
        for i in range(len(mats)):
            assert npv[i] == exp[i].np
            assert b.rows(i).equals(exp[i])


def _all_zero_matrix(p, seed):
    rng = np.random.default_rng(seed)
    nc, ncl = 80, 30
    piv = []
    for c in range(ncl):
        r = np.zeros(nc, np.int64)
        r[c] = 1
        nz = rng.choice(np.arange(c + 1, nc), size=6, replace=False)
        r[nz] = rng.integers(1, p, size=6)
        piv.append(r)
    red = []
    for k in range(25):
        coef = rng.integers(0, p, size=ncl)
        coef[k % ncl] = 1
        coef[: k % ncl] = 0
        red.append(sum(int(coef[j]) * piv[j] for j in range(ncl)) % p)
    return sort_rows(from_dense_rows(piv + red, nc, ncl, p))


@pytest.mark.parametrize("form", FORMS)
def test_batch_edge_cases(gpu_lib, form):
    mats = [
        synth.f4_shaped(60, 90, 60, 2147483629, seed=1, piv_len=(2, 10), piv_mu=np.log(5)),  # no rows to reduce
        _all_zero_matrix(32003, 2),  # every remainder is zero
        synth.f4_shaped(400, 600, 240, 65521, seed=5, piv_len=(4, 30), piv_mu=np.log(10), red_nnz=18),
        synth.f4_shaped(90, 2500, 50, 251, seed=6, piv_len=(4, 40), piv_mu=np.log(12), red_nnz=30),
    ]
    exp = [oracle.la_sparse_dense(m) for m in mats]
    b = gpu_lib.batch(mats)
    for _ in range(2):
        npv = b.la(form)
        for i in range(len(mats)):
            assert npv[i] == exp[i].np
            assert b.rows(i).equals(exp[i])
    assert npv[0] == 0 and npv[1] == 0


def test_batch_64_uneven_31bit(gpu_lib):
    """64 F4-shaped matrices (rows lognormal in [200, 3000], columns 1.5 x
    rows, density 0.5-2 %), each modulo its own prime below 2^31: both
    forms against the oracle on every matrix."""
    mats = synth.batch_31bit(64, seed=11, rows=(200, 3000))
    assert len({m.p for m in mats}) == 64
    b = gpu_lib.batch(mats)
    npm = b.la(FORM_MFMA)
    got_m = [b.rows(i) for i in range(len(mats))]
    npp = b.la(FORM_PLAIN)
    assert np.array_equal(npm, npp)
    for i, m in enumerate(mats):
        exp = oracle.la_sparse_dense(m, nthreads=8)
        assert npm[i] == exp.np
        assert got_m[i].equals(exp)
        assert b.rows(i).equals(exp)


@pytest.mark.parametrize("form", FORMS)
def test_batch_eager_launches_equal_the_graph(gpu_lib, form, monkeypatch):
    """GBGPU_BATCH_EAGER=1 issues the batch's launches straight on the
    stream (for tracing tools): the same rows as the graph replay.  Every
    launch is checked against the device limits on the host first; the graph
    holds one node per launch."""
    primes = [2147483647, 65521, 251, 32003]
    mats = _small_batch(primes, 21)
    exp = [oracle.la_sparse_dense(m) for m in mats]
    b = gpu_lib.batch(mats)
    b.la(form)
    info = b.info()
    assert info["graph_nodes"] > 0
    b.free()
    monkeypatch.setenv("GBGPU_BATCH_EAGER", "1")
    e = gpu_lib.batch(mats)
    for _ in range(2):
        npv = e.la(form)
        for i in range(len(mats)):
            assert npv[i] == exp[i].np
            assert e.rows(i).equals(exp[i])
    assert e.info()["graph_nodes"] == 0
    e.free()


TILES = [(64, 64), (256, 64), (512, 64), (256, 128), (512, 128)]
_TILE_CASES = {}


def _tile_case(primes):
    """Three matrices whose D' blocks span several update strips and row
    blocks, and the oracle's results (computed once per prime set)."""
    key = tuple(primes)
    if key not in _TILE_CASES:
        mats = [synth.f4_shaped(nr, int(nr * 1.5), int(nr * 0.6), p, seed=40 + k,
                                piv_len=(3, 40), piv_mu=np.log(12), red_nnz=20)
                for k, (nr, p) in enumerate(zip((1500, 2200, 3000), primes))]
        _TILE_CASES[key] = (mats, [oracle.la_sparse_dense(m) for m in mats])
    return _TILE_CASES[key]


@pytest.mark.parametrize("tile", TILES, ids=[f"{c}x{r}" for c, r in TILES])
@pytest.mark.parametrize("primes", [[2147483647, 16777213, 2147483587], [65521, 32003, 40009],
                                    [251, 241, 239]],
                         ids=["four_digits", "two_digits", "one_digit"])
def test_batch_update_tiles(gpu_lib, tile, primes, monkeypatch):
    """Every dense update tile the batch's tuner chooses from (forced with
    GBGPU_BATCH_TILE) gives the oracle's rows in the MFMA form, on blocks
    wide and tall enough for several strips and row blocks; tiles of 128
    rows exist for at most two digits and are refused above."""
    two_digits = max(primes) <= 65521
    monkeypatch.setenv("GBGPU_BATCH_TILE", f"{tile[0]},{tile[1]}")
    if tile[1] == 128 and not two_digits:
        mats = [synth.f4_shaped(300, 450, 180, p, seed=7, piv_len=(3, 30), red_nnz=12)
                for p in primes]
        b = gpu_lib.batch(mats)
        with pytest.raises(GpuError, match="GBGPU_BATCH_TILE"):
            b.la(FORM_MFMA)
        b.free()
        return
    mats, exp = _tile_case(primes)
    b = gpu_lib.batch(mats)
    for _ in range(2):
        npv = b.la(FORM_MFMA)
        for i in range(len(mats)):
            assert npv[i] == exp[i].np
            assert b.rows(i).equals(exp[i])
    assert b.info()["dense_tile"] == list(tile)
    b.free()
