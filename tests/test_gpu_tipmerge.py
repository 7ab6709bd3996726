"""tipMerge on the GPU (dense bit-row form on the matrix cores, sparse LDS
bitmap form) against the oracle's restatement of cu/tipMerge.cu, bit-exact
labels; the resident-cells form against the file form."""
import numpy as np
import pytest

from casphy_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["dense", "sparse"])
def form(request, monkeypatch):
    monkeypatch.setenv("CASPHY_TIPMERGE", request.param)
    return request.param


ALL_P = (0.0, 0.3, 0.5, 0.8, 1.0)


# list lengths: nearly all tips (1,200 at radius 1), about 500 of 1,500 at
# radius 12, about 370 of 3,001 at radius 24
@pytest.mark.parametrize("n,r,ps", [(1, 1, ALL_P), (2, 1, ALL_P), (5, 1, ALL_P), (300, 2, ALL_P),
                                    (1200, 1, (0.3, 0.8)), (1500, 12, (0.0, 0.5, 0.8)), (3001, 24, (0.8,))])
def test_tip_merge_radius_lists(hip, orc, form, n, r, ps):
    mat = synth.binary_matrix(seed=n, n=n, nsite=512, p=0.01)
    counts, idx = hip.radius_neighbor(np.ascontiguousarray(mat.T), r)
    for p in ps:
        assert np.array_equal(hip.tip_merge(counts, idx, p), orc.tip_merge(counts, idx, p)), (n, r, p)


def test_tip_merge_random_lists(hip, orc, form):
    """Asymmetric random lists of mixed lengths (the counts need no symmetry)."""
    rng = np.random.default_rng(12)
    n = 700
    lists = [np.unique(rng.integers(0, n, rng.integers(0, 90))) for _ in range(n)]
    nnb = np.array([l.size for l in lists], np.uint32)
    idx = np.concatenate(lists).astype(np.uint32)
    for p in (0.05, 0.2, 0.6):
        assert np.array_equal(hip.tip_merge(nnb, idx, p), orc.tip_merge(nnb, idx, p)), p


def test_tip_merge_window_order(hip, orc, monkeypatch):
    """More than 131,072 tips: pairs fold window by window (sparse form)."""
    monkeypatch.setenv("CASPHY_TIPMERGE", "sparse")
    n = 131072 + 5000
    rng = np.random.default_rng(21)
    perm = rng.permutation(n)
    lists, k = [None] * n, 0
    while k < n:
        c = np.sort(perm[k:k + int(rng.integers(1, 4))]).astype(np.uint32)
        for x in c:
            lists[x] = c
        k += c.size
    nnb = np.array([l.size for l in lists], np.uint32)
    idx = np.concatenate(lists)
    assert np.array_equal(hip.tip_merge(nnb, idx, 0.5), orc.tip_merge(nnb, idx, 0.5))


def test_cells_tip_merge(hip, orc):
    """On resident cells the lists and the radius mask stay on the device."""
    mat = synth.binary_matrix(seed=41, n=2000, nsite=1024, p=0.005)
    cells = hip.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
    for r in (15, 30, 60):
        cells.radius(r)
        counts, idx = cells.fetch()
        for p in (0.2, 0.8):
            got = cells.tip_merge(p)
            assert np.array_equal(got, hip.tip_merge(counts, idx, p))
            assert np.array_equal(got, orc.tip_merge(counts, idx, p))
    # after a device autoSplit round
    q = synth.query_rows(5, 2000, 100)
    radius, total = cells.autosplit(np.arange(1, 17), q, 3.0)
    counts, idx = cells.fetch()
    assert counts.sum() == total
    assert np.array_equal(cells.tip_merge(0.8), orc.tip_merge(counts, idx, 0.8))
    cells.close()
