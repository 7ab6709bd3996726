"""Trees from Newick text and transfer bootstrap support: the oracle's
reader and TBE formula pinned (hand-worked trees, round trips through a
Newick writer against the edge matrices ape's conventions give, the TBE
closed forms), and the library's host-side reader and TBE held to them;
the reference's own bootstrap trees (examples/test.bstrees) when present."""
import os

import numpy as np
import pytest

from casphy_amd import synth

REF_TREES = "/root/reference/examples/test.bstrees"


def test_read_newick_known_answer(orc):
    # tips by order of appearance (A..E = 1..5), internal nodes by '(' (root 6,
    # then (A,B) 7, then (C,D'q) 8), edges in preorder; lengths, the internal
    # label x and the comment are dropped, '' in a quoted label is a quote
    (edge, nnode, ntip, labels), = orc.read_newick("((A:1,B:2)x:3,(C,'D''q')[c]:1,E);")
    assert ntip == 5 and nnode == 3 and labels == ["A", "B", "C", "D'q", "E"]
    assert edge.tolist() == [[6, 7], [7, 1], [7, 2], [6, 8], [8, 3], [8, 4], [6, 5]]
    # numbered by a given tip order
    (edge, _, _, labels), = orc.read_newick("(A,(B,C));", tip_labels=["C", "A", "B"])
    assert labels == ["C", "A", "B"] and edge.tolist() == [[4, 2], [4, 5], [5, 3], [5, 1]]


@pytest.mark.parametrize("ntip", [2, 3, 17, 500])
def test_read_newick_round_trip(orc, ntip):
    """write(tree) read back gives the edge matrix ape's numbering gives the
    tree (synth.ape_edges, tips numbered by their labels' order), with
    multifurcations, lengths and internal labels."""
    rng = np.random.default_rng(ntip)
    parent, leaves = synth.yule_parent(ntip, rng)
    want, nnode, _ = synth.ape_edges(parent, leaves)
    text = synth.newick(parent, leaves, rng=rng)
    (edge, nn, nt, labels), = orc.read_newick(text, tip_labels=[f"t{t}" for t in range(ntip)])
    assert nt == ntip and nn == nnode and np.array_equal(edge, want)
    # a star and a caterpillar
    star = np.array([-1] + [0] * ntip)
    leaves2 = np.arange(1, ntip + 1)
    want, nnode, _ = synth.ape_edges(star, leaves2)
    (edge, nn, _, _), = orc.read_newick(synth.newick(star, leaves2), tip_labels=[f"t{t}" for t in range(ntip)])
    assert nn == 1 and np.array_equal(edge, want)


def test_tbe_closed_forms(orc):
    ysize = np.array([2, 5, 3, 1, 9])
    # identical trees: every split found (index 0): support 1; p < 2: NaN
    s = orc.tbe_support(np.zeros((4, 5), np.uint32), ysize, 10)
    assert np.array_equal(s[:3], [1.0, 1.0, 1.0]) and np.isnan(s[3]) and np.isnan(s[4])
    # indices capped at p - 1: a split never found has support 0
    s = orc.tbe_support(np.full((3, 5), 1000, np.uint32), ysize, 10)
    assert np.array_equal(s[:3], [0.0, 0.0, 0.0])
    # worked: p = 5, indices 3 and 0 -> 1 - 3 / 8
    s = orc.tbe_support(np.array([[0, 3, 1, 0, 0], [2, 0, 5, 0, 0]]), ysize, 10)
    assert s[0] == 0.5 and s[1] == 0.625 and s[2] == 0.25


def product():
    from casphy_amd import api
    return api


def test_library_reader_and_tbe_match_oracle(orc):
    api = product()
    rng = np.random.default_rng(3)
    texts = ["((A:1,B:2)x:3,(C,'D''q')[c]:1,E);\n(A,(B,C),(E,'D''q'));", "A;", "(,);", "((A)B,C);"]
    for ntip in (2, 50, 3000):
        parent, leaves = synth.yule_parent(ntip, rng)
        texts.append(" [tree]\n" + synth.newick(parent, leaves, rng=rng) + "\n" + synth.newick(parent, leaves))
    for text in texts:
        got, want = api.read_newick(text), orc.read_newick(text)
        assert len(got) == len(want)
        for g, w in zip(got, want):
            assert np.array_equal(g[0], w[0]) and g[1:] == w[1:], text[:40]
    ym = rng.integers(0, 40, (7, 300), dtype=np.uint32)
    ys = rng.integers(1, 500, 300, dtype=np.uint32)
    a, b = api.tbe_support(ym, ys, 500), orc.tbe_support(ym, ys, 500)
    assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


@pytest.mark.parametrize("text", ["(A,B", "(A,B));", "(A,B)", "((A,B),C;", "(A,B),C);", "(A,[x B);"])
def test_library_reader_rejects_malformed(text):
    api = product()
    with pytest.raises(api.CasphyError) as e:
        api.read_newick(text)
    assert "newick" in str(e.value)


def test_library_reader_tip_order_checked():
    api = product()
    with pytest.raises(api.CasphyError, match="not in the tip order"):
        api.read_newick("(A,(B,C));", tip_labels=["A", "B", "D"])
    with pytest.raises(api.CasphyError, match="appears twice"):
        api.read_newick("(A,(B,A));", tip_labels=["A", "B", "C"])


@pytest.mark.skipif(not os.path.exists(REF_TREES), reason="reference example trees not present")
def test_reference_bootstrap_trees():
    """examples/test.bstrees: 11 trees of 100,000 tips, multifurcating; the
    library and the oracle read the first identically, every tree has the
    same tips, and each is a valid cladewise edge matrix (the clade planner
    accepts it)."""
    api = product()
    from oracle import oracle
    text = open(REF_TREES, "rb").read()
    trees = api.read_newick(text)
    assert len(trees) == 11
    edge, nnode, ntip, labels = trees[0]
    assert ntip == 100000 and nnode == 99389 and edge.shape == (ntip + nnode - 1, 2)
    first = text[: text.index(b";") + 1]
    (e2, n2, t2, l2), = oracle.read_newick(first)
    assert np.array_equal(edge, e2) and (n2, t2, l2) == (nnode, ntip, labels)
    names = set(labels)
    for e, nn, nt, lab in trees[1:]:
        assert nt == ntip and set(lab) == names and e.shape == (nt + nn - 1, 2)
    aligned = api.read_newick(text, tip_labels=labels)
    assert all(t[3] == labels for t in aligned)
    for e, nn, nt, _ in aligned[:2]:  # children of each internal node leave it after its own edge
        seen = np.zeros(nt + nn + 2, bool)
        seen[nt + 1] = True
        for a, d in e:
            assert seen[a]
            seen[d] = True


def test_native_spr_generator(orc):
    """The native bootstrap-tree generator: no moves gives ape_edges of the
    tree itself; with moves, every replicate is a valid cladewise binary tree
    on the same tips (the oracle's nodeDesBig accepts it), replicates differ
    from the reference and from each other, and a seed repeats its trees."""
    rng = np.random.default_rng(30)
    ntip = 3000
    parent, leaves = synth.yule_parent(ntip, rng)
    e0, n0, _ = synth.ape_edges(parent, leaves)
    (z, nz), = synth.spr_edges_batch(parent, leaves, 0, 1, 7)
    assert nz == n0 and np.array_equal(z, e0)
    reps = synth.spr_edges_batch(parent, leaves, 150, 4, 7)
    again = synth.spr_edges_batch(parent, leaves, 150, 4, 7)
    Y, _ = orc.clade_bitsets(e0, n0, ntip)
    seen = []
    for (e, n), (e2, n2) in zip(reps, again):
        assert n == ntip - 1 and e.shape == (2 * ntip - 2, 2)
        assert np.array_equal(e, e2) and n == n2
        assert sorted(e[e[:, 1] <= ntip, 1].tolist()) == list(range(1, ntip + 1))
        X, sizes = orc.clade_bitsets(e, n, ntip)
        assert X.shape == Y.shape and np.all(sizes >= 2)
        ymin, _, _, _ = orc.transfer(Y, X, ntip)
        assert ymin.max() > 0  # some reference clade was broken
        seen.append(X.tobytes())
    assert len(set(seen)) == len(seen)
