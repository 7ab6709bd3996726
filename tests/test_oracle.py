"""Pins the CPU oracle (oracle/casphy_oracle.c) against independent checks.

CPU only.  Each oracle function is compared with a restatement that works
differently (closed forms on unpacked bits, IUPAC sets on text, O(n^3)
lexicographic joins on 3-valued states, R's reOrderEdge2 + factor relabel,
numpy Hamming), plus invariants and special cases.
"""
import itertools

import numpy as np
import pytest

from casphy_amd import synth
from tests import brute


def rand_bits(rng, n, W, p=0.1):
    return synth.pack_bits(rng.random((n, 32 * W)) < p)


# ------------------------------------------------------------------ scores
def test_score_bit_closed_form(orc):
    rng = np.random.default_rng(0)
    mat = rand_bits(rng, 40, 3, 0.2)
    S = orc.score_bit_matrix(mat)
    ref = brute.score_matrix_bits(mat)
    assert np.array_equal(S.astype(np.int64), ref)
    assert np.array_equal(S, S.T)
    assert np.array_equal(np.diag(S), synth.unpack_bits(mat).sum(1))


def test_score_bit_integer_weights(orc):
    rng = np.random.default_rng(1)
    mat = rand_bits(rng, 24, 2, 0.3)
    w = rng.integers(0, 9, 64).astype(np.float32)
    S = orc.score_bit_matrix(mat, w)
    assert np.array_equal(S.astype(np.int64), brute.score_matrix_bits(mat, w.astype(np.int64)))


def test_score_bit_zero_rows(orc):
    mat = np.zeros((3, 2), np.uint32)
    mat[1, 0] = 0xFFFFFFFF
    S = orc.score_bit_matrix(mat)
    assert S[0].sum() == 0 and S[1, 1] == 32 and S[1, 2] == 0


def test_score_indel_onehot(orc):
    mat = synth.indel_matrix(seed=5, depth=4, nsite=30, p=0.1, nallele=6)
    w = np.ones((mat.max(), mat.shape[1]), dtype=np.float32)
    ref = brute.score_matrix_indel(mat)
    for i in range(mat.shape[0]):
        for j in range(mat.shape[0]):
            assert orc.score_indel(mat[i], mat[j], w) == ref[i, j]
    rng = np.random.default_rng(2)
    wi = rng.integers(0, 5, w.shape).astype(np.float32)
    refw = brute.score_matrix_indel(mat, wi.astype(np.int64))
    assert orc.score_indel(mat[3], mat[7], wi) == refw[3, 7]


def test_score_indel_matrix_closed_form(orc):
    """The all-pairs form (diagonal included) equals the one-hot closed form,
    unit and integer weights."""
    mat = synth.indel_matrix(seed=6, depth=6, nsite=40, p=0.08, nallele=9)
    w = np.ones((mat.max(), mat.shape[1]), dtype=np.float32)
    assert np.array_equal(orc.score_indel_matrix(mat, w), brute.score_matrix_indel(mat))
    wi = np.random.default_rng(3).integers(0, 6, w.shape)
    assert np.array_equal(orc.score_indel_matrix(mat, wi.astype(np.float32)),
                          brute.score_matrix_indel(mat, wi.astype(np.int64)))


def test_score_dna_iupac_text(orc):
    rng = np.random.default_rng(3)
    alphabet = "ATCGUWMRYKSHDVBN-.X"
    code = {c: 0x10 for c in alphabet}
    code.update({"A": 1, "T": 2, "U": 2, "C": 4, "G": 8, "W": 3, "M": 5, "R": 9, "Y": 6, "K": 10, "S": 12,
                 "H": 7, "D": 11, "V": 13, "B": 14, "N": 15})
    for L in (1, 7, 33, 64):
        s1 = "".join(rng.choice(list(alphabet), L))
        s2 = "".join(rng.choice(list(alphabet), L))
        enc = synth.pack_bytes4(np.array([[code[c] for c in s1], [code[c] for c in s2]], dtype=np.uint8))
        assert orc.score_dna(enc[0], enc[1]) == brute.dna_similarity_text(s1, s2)


# ------------------------------------------------------------------ radius
@pytest.mark.parametrize("r", [0, 1, 3, 6, 200])
def test_radius_bit_brute(orc, r):
    rng = np.random.default_rng(4)
    mat = rand_bits(rng, 70, 2, 0.08)
    counts, idx = orc.radius_bit(mat, r)
    S = brute.score_matrix_bits(mat)
    c2, i2 = brute.radius_lists_from_matrix(S, r)
    assert np.array_equal(counts, c2) and np.array_equal(idx, i2)


def test_radius_monotone_and_self(orc):
    rng = np.random.default_rng(5)
    mat = rand_bits(rng, 50, 1, 0.2)
    prev = None
    pop = synth.unpack_bits(mat).sum(1)
    for r in range(0, 12):
        counts, idx = orc.radius_bit(mat, r)
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        for i in range(50):
            li = idx[off[i]:off[i + 1]]
            assert np.all(np.diff(li.astype(np.int64)) > 0)
            assert (i in li) == (pop[i] >= r)
        if prev is not None:
            assert np.all(counts <= prev)
        prev = counts


def test_radius_dna_and_try(orc):
    mat = synth.dna_matrix(seed=6, n=60, ncol=90)
    S = np.array([[orc.score_dna(mat[i], mat[j]) for j in range(60)] for i in range(60)])
    assert np.array_equal(S, S.T)
    for r in (40, 70, 85):
        counts, idx = orc.radius_dna(mat, r)
        c2, i2 = brute.radius_lists_from_matrix(S, r)
        assert np.array_equal(counts, c2) and np.array_equal(idx, i2)
        q = np.array([0, 5, 5, 59], np.uint32)
        assert np.array_equal(orc.try_radius_dna(mat, r, q), c2[q])


def test_try_radius_bit(orc):
    rng = np.random.default_rng(7)
    mat = rand_bits(rng, 80, 2, 0.1)
    S = brute.score_matrix_bits(mat)
    q = rng.integers(0, 80, 33).astype(np.uint32)
    for r in (1, 4, 9):
        assert np.array_equal(orc.try_radius_bit(mat, r, q), (S[q] >= r).sum(1))


# ------------------------------------------------------------------ ancestral rules
@pytest.mark.parametrize("anc", [0, 1])
def test_ancestral_rule_all_pairs_via_join(orc, anc):
    """4 tips built so that the joins are ((0,1),(2,3)); site values at the
    two internal nodes cover {0,1,?}^2 and the root state is checked."""
    enc = brute.ENCODE[anc]
    vals = [0, 1, brute.Q]
    tips_for = {0: (0, 0), 1: (1, 1), brute.Q: (1, 0)}
    sites = list(itertools.product(vals, vals))
    W = 1
    mat = np.zeros((4, W), np.uint32)
    for k, (u, v) in enumerate(sites):
        a, b = tips_for[u]
        c, d = tips_for[v]
        for t, bit in zip(range(4), (a, b, c, d)):
            mat[t, 0] |= np.uint32(bit) << np.uint32(k)
    # anchor bits: 20 bits shared by tips 0,1 and 14 by tips 2,3 make the join
    # order (0,1), (2,3), root regardless of the test sites
    for k in range(10, 30):
        mat[0, 0] |= np.uint32(1) << np.uint32(k)
        mat[1, 0] |= np.uint32(1) << np.uint32(k)
    for k in range(30, 32):
        mat[2, 0] |= np.uint32(1) << np.uint32(k)
        mat[3, 0] |= np.uint32(1) << np.uint32(k)
    edge, ns, nb = orc.join_bit(mat, ancestral_type=anc)
    assert [tuple(e) for e in edge[:2]] == [(4, 0), (4, 1)]
    assert [tuple(e) for e in edge[2:4]] == [(5, 2), (5, 3)]
    for k, (u, v) in enumerate(sites):
        for node, val in ((4, u), (5, v)):
            assert ((ns[node, 0] >> k) & 1) == (1 if val == 1 else 0)
            assert ((nb[node, 0] >> k) & 1) == enc[val][1]
        want = brute.RULE[anc][(u, v)]
        assert ((ns[6, 0] >> k) & 1) == (1 if want == 1 else 0), (u, v)
        assert ((nb[6, 0] >> k) & 1) == enc[want][1], (u, v)


# ------------------------------------------------------------------ joins
@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("anc", [0, 1])
def test_join_bit_vs_lexicographic(orc, seed, anc):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(2, 24))
    mat = rand_bits(rng, n, 1, float(rng.choice([0.05, 0.2, 0.5])))
    edges, state = brute.bin_join(mat, anc)
    e1, ns1, _ = orc.join_bit(mat, ancestral_type=anc)
    e2, ns2, _ = orc.join_bit(mat, ancestral_type=anc, lazy=True)
    assert np.array_equal(e1, edges)
    assert np.array_equal(e2, edges)
    assert np.array_equal(synth.unpack_bits(ns1), state)
    assert np.array_equal(ns1, ns2)


def test_join_bit_weighted_vs_lexicographic(orc):
    rng = np.random.default_rng(11)
    mat = rand_bits(rng, 14, 1, 0.3)
    w = rng.integers(0, 4, 32)
    edges, _ = brute.bin_join(mat, 1, w)
    e1, _, _ = orc.join_bit(mat, w.astype(np.float32), 1)
    e2, _, _ = orc.join_bit(mat, w.astype(np.float32), 1, lazy=True)
    assert np.array_equal(e1, edges) and np.array_equal(e2, edges)


def test_join_bit_invariants(orc):
    mat = synth.binary_matrix(seed=12, n=300, nsite=256, p=0.02)
    e1, ns1, _ = orc.join_bit(mat)
    e2, ns2, _ = orc.join_bit(mat, lazy=True)
    n = 300
    assert np.array_equal(e1, e2) and np.array_equal(ns1, ns2)
    child = e1[:, 1]
    assert sorted(child.tolist()) == list(range(2 * n - 2))
    assert np.all(e1[0::2, 0] == np.arange(n, 2 * n - 1)) and np.all(e1[1::2, 0] == e1[0::2, 0])
    assert np.all(e1[0::2, 1] < e1[1::2, 1])  # first child is the smaller id


@pytest.mark.parametrize("seed", range(5))
def test_join_indel_vs_lexicographic(orc, seed):
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(2, 20))
    mat = rng.integers(-1, 4, (n, 12)).astype(np.int32)
    w = np.ones((max(1, mat.max()), 12))
    ts = rng.integers(1, 5, n).astype(np.uint32)
    for min_size, ratio in ((0, 2.0), (3, 1.5)):
        edges, states = brute.indel_join(mat, None, ts, min_size, ratio)
        e1, ns1, br1 = orc.join_indel(mat, w, ts, min_size, ratio)
        e2, ns2, br2 = orc.join_indel(mat, w, ts, min_size, ratio, lazy=True)
        assert np.array_equal(e1, edges) and np.array_equal(e2, edges)
        assert np.array_equal(ns1, states) and np.array_equal(ns2, states)
        br = [int(((states[p] == 0) & (states[c] > 0)).sum()) for p, c in edges]
        assert np.array_equal(br1, br)


def test_join_indel_weighted(orc):
    rng = np.random.default_rng(13)
    mat = rng.integers(-1, 3, (12, 10)).astype(np.int32)
    w = rng.integers(0, 4, (2, 10))
    edges, _ = brute.indel_join(mat, w)
    e1, _, _ = orc.join_indel(mat, w.astype(np.float32))
    assert np.array_equal(e1, edges)


# ------------------------------------------------------------------ emission
@pytest.mark.parametrize("seed", range(4))
def test_cladewise_vs_R(orc, seed):
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(2, 60))
    mat = rand_bits(rng, n, 1, 0.2)
    edge, _, _ = orc.join_bit(mat)
    eo, io = orc.cladewise(edge, n)
    e2, io2 = brute.cladewise_R(edge, n)
    assert np.array_equal(eo, e2) and np.array_equal(io, io2)
    assert eo[0, 0] == n and io[0] == 2 * n - 2
    # preorder: every internal parent appears (as a child) before its own edges
    seen = {n}
    for p, c in eo:
        assert p in seen
        seen.add(int(c))


# ------------------------------------------------------------------ transfer
def test_transfer_brute(orc):
    rng = np.random.default_rng(14)
    ntip = 75
    W = 3
    def bits(m):
        b = rng.random((m, ntip)) < 0.4
        return synth.pack_bits(np.pad(b, ((0, 0), (0, 32 * W - ntip))))
    Y, X = bits(20), bits(33)
    got = orc.transfer(Y, X, ntip)
    want = brute.transfer_brute(Y, X, ntip)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_transfer_identical_trees_and_complement(orc):
    rng = np.random.default_rng(15)
    parent, leaves = synth.yule_parent(40, rng)
    B = synth.clade_bitsets(parent, leaves, 40)
    ymin, ysize, xmin, xsize = orc.transfer(B, B, 40)
    assert np.all(ymin == 0) and np.all(xmin == 0)
    comp = synth.pack_bits(~synth.unpack_bits(B, 40))
    ymin2, ysize2, _, _ = orc.transfer(B, comp, 40)
    assert np.all(ymin2 == 0) and np.array_equal(ysize2, ysize)


# ------------------------------------------------------------------ tipMerge
def _lists_arrays(lists):
    nnb = np.array([len(l) for l in lists], dtype=np.uint32)
    idx = np.concatenate([np.asarray(sorted(l), np.uint32) for l in lists]) if lists else np.zeros(0, np.uint32)
    return nnb, idx


def test_list_shared_vs_numpy(orc):
    rng = np.random.default_rng(3)
    for _ in range(50):
        a = np.unique(rng.integers(0, 500, rng.integers(0, 120)))
        b = np.unique(rng.integers(0, 500, rng.integers(0, 120)))
        assert orc.list_shared(a, b) == np.intersect1d(a, b).size


def _clique_lists(n, rng, max_size=4):
    """Tips shuffled into disjoint cliques; each tip's list is its clique."""
    perm = rng.permutation(n)
    cliques, k = [], 0
    while k < n:
        s = int(rng.integers(1, max_size + 1))
        cliques.append(sorted(int(x) for x in perm[k:k + s]))
        k += s
    lists = [None] * n
    for c in cliques:
        for x in c:
            lists[x] = c
    return cliques, lists


def _clique_labels(n, cliques, window=131072):
    """Closed form for disjoint cliques: a clique of two or more tips becomes a
    category when its first pair is folded -- in the window of its second
    smallest tip, at its smallest tip -- so categories are numbered by
    (second smallest // window, smallest); single tips follow in tip order."""
    multi = sorted((c for c in cliques if len(c) > 1), key=lambda c: (c[1] // window, c[0]))
    out = np.full(n, -1, np.int64)
    for lab, c in enumerate(multi):
        out[c] = lab
    nxt = len(multi)
    for i in range(n):
        if out[i] < 0:
            out[i] = nxt
            nxt += 1
    return out


@pytest.mark.parametrize("n", [1, 2, 97, 2000])
def test_tip_merge_disjoint_cliques(orc, n):
    rng = np.random.default_rng(n)
    cliques, lists = _clique_lists(n, rng)
    nnb, idx = _lists_arrays(lists)
    for p in (0.0, 0.5, 0.99):
        assert np.array_equal(orc.tip_merge(nnb, idx, p), _clique_labels(n, cliques))
    # p = 1: s > |N| p never holds (s <= |N|), so no pair joins
    assert np.array_equal(orc.tip_merge(nnb, idx, 1.0), np.arange(n))


def test_tip_merge_window_order(orc):
    """Past 131,072 tips the pairs are folded window by window: cliques whose
    second tip lies in the second window are numbered after every clique of
    the first window, whatever their smallest tip."""
    n = 131072 + 3000
    rng = np.random.default_rng(8)
    cliques, lists = _clique_lists(n, rng, max_size=3)
    nnb, idx = _lists_arrays(lists)
    got = orc.tip_merge(nnb, idx, 0.5)
    assert np.array_equal(got, _clique_labels(n, cliques))
    assert not np.array_equal(got, _clique_labels(n, cliques, window=n))  # the window matters here


def test_tip_merge_either_fraction(orc):
    """s > |N(j)| p OR s > |N(t)| p, strict: tips 10 < 11 with |N| = 10 and 4."""
    def run(l10, l11, p):
        lists = [[k] for k in range(10)] + [l10, l11]
        nnb, idx = _lists_arrays(lists)
        return orc.tip_merge(nnb, idx, p)
    big, small = list(range(8)) + [10, 11], [0, 8, 9, 11]      # s = |{0, 11}| = 2
    joined = np.array(list(range(1, 11)) + [0, 0])
    alone = np.arange(12)
    # second operand (|N(t)| = 4): 2 > 4p iff p < 0.5
    assert np.array_equal(run(big, small, 0.49), joined)
    assert np.array_equal(run(big, small, 0.5), alone)
    # first operand (|N(j)| = 4): s = |{0, 10, 11}| = 3 > 4p iff p < 0.75
    small2, big2 = [0, 8, 10, 11], list(range(8)) + [10, 11]
    assert np.array_equal(run(small2, big2, 0.74), joined)
    assert np.array_equal(run(small2, big2, 0.75), alone)


def test_tip_merge_labels_follow_links(orc):
    """On radius lists: tips with one label are connected through the lists,
    and lowering p never splits a cluster."""
    mat = synth.binary_matrix(seed=31, n=400, nsite=256, p=0.02)
    counts, idx = orc.radius_bit(mat, 3)
    off = np.zeros(401, np.int64)
    off[1:] = np.cumsum(counts)
    # connected components of the list graph (union-find over all links)
    par = np.arange(400)

    def find(x):
        while par[x] != x:
            par[x] = par[par[x]]
            x = par[x]
        return x
    for i in range(400):
        for t in idx[off[i]:off[i + 1]]:
            a, b = find(i), find(int(t))
            if a != b:
                par[a] = b
    comp = np.array([find(i) for i in range(400)])
    prev = None
    for p in (0.9, 0.6, 0.3, 0.0):
        lab = orc.tip_merge(counts, idx, p)
        for v in np.unique(lab):
            assert np.unique(comp[lab == v]).size == 1
        if prev is not None:  # clusters at a higher p are inside one cluster at a lower p
            for v in np.unique(prev):
                assert np.unique(lab[prev == v]).size == 1
        prev = lab


# ------------------------------------------------------------------ nodeKeep
def _joined_tree_bit(orc, seed, n=200, nsite=256):
    mat = synth.binary_matrix(seed=seed, n=n, nsite=nsite, p=0.02)
    edge, ns, _ = orc.join_bit(mat)
    eo, order = orc.cladewise(edge, n)
    return mat, eo, ns[order]


def _keep_by_tree_walk(eo, retain, nnode):
    """Kept set of the skip rule on a tree whose internal ids are in preorder:
    a node is kept when it passes and no proper ancestor is kept."""
    ntip = nnode + 1
    parent = {int(d): int(a) for a, d in eo}
    kept = []
    under_kept = {}
    for v in sorted(d for d in parent if d >= ntip):
        p = parent[v]
        blocked = under_kept.get(p, False) or (p in kept)
        under_kept[v] = blocked
        if not blocked and retain[v - ntip]:
            kept.append(v)
    return kept


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_node_keep_bit_tree_walk(orc, seed):
    mat, eo, ns2 = _joined_tree_bit(orc, 60 + seed)
    nnode = ns2.shape[0]
    assert all(d > a for a, d in eo if d > nnode)  # internal ids in preorder
    rng = np.random.default_rng(seed)
    bits = synth.unpack_bits(ns2).astype(bool)
    for nout in (0, 1, 40):
        out = mat[rng.choice(mat.shape[0], nout, replace=False)] if nout else np.zeros((0, ns2.shape[1]), np.uint32)
        ob = synth.unpack_bits(out).astype(bool) if nout else np.zeros((0, bits.shape[1]), bool)
        missing = (bits[:, None, :] & ~ob[None, :, :]).sum(2) if nout else np.zeros((nnode, 0), int)
        for cutoff in (0, 1, 2, 5):
            retain = np.all(missing >= cutoff, axis=1)
            want = _keep_by_tree_walk(eo, retain, nnode)
            assert orc.node_keep(eo, ns2, out, cutoff).tolist() == want, (nout, cutoff)


def test_node_keep_indel_tree_walk(orc):
    rng = np.random.default_rng(9)
    mat = synth.indel_matrix(seed=5, depth=7, nsite=40)
    n, ncol = mat.shape
    w = np.ones((mat.max(), ncol), dtype=np.float32)
    edge, ns, _ = orc.join_indel(mat, w)
    eo, order = orc.cladewise(edge, n)
    ns2 = (ns[order] + 1).astype(np.uint32)          # .nodeState2 holds states + 1
    nnode = ns2.shape[0]
    out = (mat[rng.choice(n, 30, replace=False)] + 1).astype(np.uint32)   # .tboutmat: outmat + 1
    a = ns2[:, None, :].astype(np.int64)
    b = out[None, :, :].astype(np.int64)
    missing = ((a > 1) & (b >= 1) & (b != a)).sum(2)
    for cutoff in (1, 2, 4):
        retain = np.all(missing >= cutoff, axis=1)
        assert orc.node_keep(eo, ns2, out, cutoff, indel=True).tolist() == _keep_by_tree_walk(eo, retain, nnode)


# ------------------------------------------------------------------ clade bitsets
def _clades_by_walk(edge, nnode, ntip):
    """Clade of every internal node by walking from each tip to the root."""
    par = {int(d): int(a) for a, d in edge}
    nbin = ((ntip - 1) >> 5) + 1
    rows = np.zeros((nnode + ntip + 1, nbin), np.uint32)
    for t in range(1, ntip + 1):
        v = par.get(t)
        while v is not None:
            rows[v, (t - 1) >> 5] |= np.uint32(1) << np.uint32((t - 1) & 31)
            v = par.get(v)
    return rows


def _random_multifurcating(rng, ntip):
    """Random rooted tree with internal nodes of 2..4 children (parent array)."""
    nodes = list(range(ntip))
    parent = {}
    nxt = ntip
    while len(nodes) > 1:
        k = min(len(nodes), int(rng.integers(2, 5)))
        pick = rng.choice(len(nodes), k, replace=False)
        for i in pick:
            parent[nodes[i]] = nxt
        nodes = [x for j, x in enumerate(nodes) if j not in set(pick.tolist())] + [nxt]
        nxt += 1
    root = nodes[0]
    # renumber so the root is 0 and tips follow their order
    order = [root] + [v for v in range(nxt) if v != root]
    idx = {v: i for i, v in enumerate(order)}
    par = np.full(nxt, -1, np.int64)
    for c, p in parent.items():
        par[idx[c]] = idx[p]
    leaves = np.array([idx[t] for t in range(ntip)], np.int64)
    return par, leaves


@pytest.mark.parametrize("ntip,multi", [(2, False), (3, False), (37, False), (500, False), (64, True), (333, True)])
def test_clade_bitsets_walk(orc, ntip, multi):
    rng = np.random.default_rng(ntip)
    parent, leaves = _random_multifurcating(rng, ntip) if multi else synth.yule_parent(ntip, rng)
    edge, nnode, nt = synth.ape_edges(parent, leaves)
    rows, sizes = orc.clade_bitsets(edge, nnode, nt)
    walk = _clades_by_walk(edge, nnode, nt)
    internal_children = [int(d) for a, d in edge[::-1] if d > nt]   # backward scan order
    assert rows.shape[0] == nnode - 1 == len(internal_children)
    for k, v in enumerate(internal_children):
        assert np.array_equal(rows[k], walk[v]), k
    pop = synth.unpack_bits(walk).sum(1)
    # the reference's return vector: internal node nnode-1-k (0-based), k < nnode-1
    assert [int(x) for x in sizes] == [int(pop[nt + 1 + (nnode - 1 - k)]) for k in range(nnode - 1)]
