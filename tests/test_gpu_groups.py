"""Group-mode autoSplit search (casphy_cells_split_groups): many groups of one
resident matrix searched at once, each group held to what autoSplit
(R/CasPhy_functions.R:1964-2082) computes on its own sub-matrix mat[rows, ]:
the oracle's tryRadius counts of the group's queries, the stopping rule with
the group's bound nNeighborBound - log10(n_g), the oracle's radius lists of
the sub-matrix at the chosen radius (ids local to the group), and tipMerge on
those lists."""
import numpy as np
import pytest

from casphy_amd import synth

pytestmark = pytest.mark.gpu


def autosplit_rule(counts, radii, bound):
    for k, r in enumerate(radii):
        if np.log10(max(counts[k].mean(), 1e-30)) <= bound:
            return int(r)
    return -1


@pytest.fixture(params=["fp4", "64x2", "fp4d", "valu"])
def tiles(request, monkeypatch):
    """The group launches on the fp4 and int8 matrix-core tiles, and the
    VALU cells that run the groups in turn."""
    monkeypatch.setenv("CASPHY_MFMA", "0" if request.param == "valu" else "1")
    if request.param == "valu":
        monkeypatch.delenv("CASPHY_MFMA_VARIANT", raising=False)
    else:
        monkeypatch.setenv("CASPHY_MFMA_VARIANT", request.param)
    return request.param


def make_groups(rng, n, sizes):
    perm = rng.permutation(n)
    out, at = [], 0
    for sz in sizes:
        out.append(np.sort(perm[at:at + sz]).astype(np.uint32))
        at += sz
    return out


def check(cells, orc, kind, mat, groups, queries, radii, nb):
    rad, nidx = cells.split_groups(groups, radii, queries, nb)
    assert rad.size == len(groups) and nidx.size == len(groups)
    try_ref = orc.try_radius_bit if kind == "bit" else orc.try_radius_dna
    lists_ref = orc.radius_bit if kind == "bit" else orc.radius_dna
    chosen = []
    for g, rows in enumerate(groups):
        sub = np.ascontiguousarray(mat[rows])
        ref = np.stack([try_ref(sub, int(r), queries[g]) for r in radii])
        want = autosplit_rule(ref, radii, nb - np.log10(len(rows)))
        assert rad[g] == want, (g, len(rows))
        chosen.append(want)
        counts, idx = cells.group_fetch(g)
        if want < 0:
            assert nidx[g] == 0 and not counts.any(), g
            continue
        c2, i2 = lists_ref(sub, want)
        assert nidx[g] == i2.size and np.array_equal(counts, c2) and np.array_equal(idx, i2), (g, len(rows))
    return chosen


def test_split_groups_bit(hip, orc, tiles):
    rng = np.random.default_rng(41)
    n = 2600
    mat = synth.binary_matrix(seed=41, n=n, nsite=256, p=0.01)
    cells = hip.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
    sizes = [1, 5, 127, 128, 129, 300, 700, 1000]
    groups = make_groups(rng, n, sizes)
    groups.append(groups[6][::3].copy())  # rows of another group: groups are independent problems
    queries = [rng.integers(0, len(g), min(len(g), 150)).astype(np.uint32) for g in groups]
    radii = np.arange(1, 17) * 2
    outcomes = []
    for nb in (3.2, 1.6, 0.6):  # the same handle again: buffers are reused
        outcomes += check(cells, orc, "bit", mat, groups, queries, radii, nb)
    assert min(outcomes) < 0 <= max(outcomes) and len(set(outcomes)) >= 3


def test_split_groups_64_radii_and_regrouping(hip, orc, tiles):
    """64 radii (four counting passes), then a second grouping of more rows."""
    rng = np.random.default_rng(43)
    n = 1800
    mat = synth.binary_matrix(seed=43, n=n, nsite=512, p=0.008)
    cells = hip.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
    radii = np.arange(1, 65)
    g1 = make_groups(rng, n, [40, 260, 500])
    q1 = [rng.integers(0, len(g), 64).astype(np.uint32) for g in g1]
    check(cells, orc, "bit", mat, g1, q1, radii, 2.0)
    g2 = make_groups(rng, n, [1, 2, 3, 300, 1494])
    q2 = [rng.integers(0, len(g), 200).astype(np.uint32) for g in g2]
    check(cells, orc, "bit", mat, g2, q2, radii, 2.5)


@pytest.mark.parametrize("form", ["mfma", "valu"])
def test_split_groups_dna(hip, orc, monkeypatch, form):
    monkeypatch.setenv("CASPHY_MFMA", "1")
    monkeypatch.delenv("CASPHY_MFMA_VARIANT", raising=False)
    monkeypatch.setenv("CASPHY_DNA_FORM", form)
    rng = np.random.default_rng(44)
    n = 1500
    mat = synth.dna_matrix(seed=44, n=n, ncol=300)
    cells = hip.Cells(mat, "dna")
    groups = make_groups(rng, n, [3, 130, 257, 400, 600])
    queries = [rng.integers(0, len(g), 100).astype(np.uint32) for g in groups]
    radii = np.array([200, 230, 250, 265, 275, 285, 292, 299])
    outcomes = []
    for nb in (4.0, 3.0):
        outcomes += check(cells, orc, "dna", mat, groups, queries, radii, nb)
    assert max(outcomes) >= 0


def test_group_tip_merge(hip, orc):
    """tipMerge on a group's lists equals the oracle's on the fetched lists,
    for every group that chose a radius (and singletons for one that did not)."""
    rng = np.random.default_rng(45)
    n = 2000
    mat = synth.binary_matrix(seed=45, n=n, nsite=512, p=0.01)
    cells = hip.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
    groups = make_groups(rng, n, [90, 400, 1500])
    queries = [rng.integers(0, len(g), 120).astype(np.uint32) for g in groups]
    rad, nidx = cells.split_groups(groups, np.arange(2, 34, 2), queries, 4.5)
    assert (rad >= 0).sum() >= 2 and (rad < 0).any()
    for g in range(len(groups)):
        counts, idx = cells.group_fetch(g)
        member = cells.group_tip_merge(g, 0.8)
        assert np.array_equal(member, orc.tip_merge(counts, idx, 0.8)), g


def test_split_groups_errors(hip):
    mat = synth.binary_matrix(seed=46, n=300, nsite=128, p=0.01)
    cells = hip.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
    g = [np.arange(100, dtype=np.uint32), np.arange(100, 300, dtype=np.uint32)]
    with pytest.raises(Exception, match="out of range"):
        cells.split_groups(g, [1, 2], [np.array([100], np.uint32), np.array([0], np.uint32)])
    with pytest.raises(Exception, match="out of range"):
        cells.split_groups([np.array([0, 300], np.uint32)], [1, 2], [np.array([0], np.uint32)])
    with pytest.raises(Exception, match="no query"):
        cells.split_groups(g, [1, 2], [np.array([1], np.uint32), np.array([], np.uint32)])
    with pytest.raises(Exception, match="ascending"):
        cells.split_groups(g, [3, 2], [np.array([1], np.uint32), np.array([1], np.uint32)])
    with pytest.raises(Exception, match="no such group"):
        cells.group_fetch(7)
