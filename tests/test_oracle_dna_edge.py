"""Pins of the oracle's DNA-format-edge restatements: voteRoot2 / findRoot
(cu/voteRoot2.cu, src/functions.cpp:4170-4410) and fasta2bin
(cu/fasta2bin.cu), each against an independent numpy formulation, known
answers worked by hand, and the reference's own example alignment
(tests/golden/run1_1024_head.fasta: the first 48 records of
examples/run1_1024.fasta)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN_FASTA = os.path.join(HERE, "golden", "run1_1024_head.fasta")


# ---------------------------------------------------------------- root votes
def np_counts(state, edge, samples):
    """Shared positions of every sample with every node's edge state: bytes
    unpacked, AND, base bits tested -- no word loop."""
    nnode, nbin = state.shape
    es = np.zeros_like(state)
    if edge.size:
        a, d = edge[:, 0].astype(np.int64) - 1, edge[:, 1].astype(np.int64) - 1
        es[d] = state[a] | state[d]
    eb = es.view(np.uint8).reshape(nnode, nbin * 4) & 0x0F
    sb = samples.view(np.uint8).reshape(samples.shape[0], nbin * 4) & 0x0F
    return ((eb[None, :, :] & sb[:, None, :]) != 0).sum(-1)


def bitrev(x, bits):
    return sum(((x >> k) & 1) << (bits - 1 - k) for k in range(bits))


def np_vote_root2(counts):
    """The winner as a closed form: among the nodes of the largest count, the
    first by (bit-reversed block of 8 mod 32, block / 32, bit-reversed node
    mod 8) -- the order in which halving-stride tournaments whose lower slot
    keeps ties give way."""
    nnode = counts.shape[1]
    v = np.arange(nnode)
    order = np.lexsort((bitrev(v % 8, 3), (v // 8) // 32, bitrev((v // 8) % 32, 5)))
    rank = np.empty(nnode, np.int64)
    rank[order] = np.arange(nnode)
    root = np.empty(counts.shape[0], np.int64)
    for s, c in enumerate(counts):
        cand = np.flatnonzero(c == c.max())
        root[s] = cand[np.argmin(rank[cand])]
    return root, counts.max(axis=1)


def random_tree_edges(nnode, rng, nroot=1):
    """A random forest on nodes 1..nnode with `nroot` roots, edges shuffled."""
    perm = rng.permutation(nnode) + 1
    edges = [(perm[rng.integers(0, i)], perm[i]) for i in range(nroot, nnode)]
    e = np.array(edges, np.uint32).reshape(-1, 2)
    return e[rng.permutation(e.shape[0])]


def tie_rich(rng, nnode, nbin, nsample, nmat=None):
    """Few distinct codes on few positions: many equal counts."""
    alphabet = np.array([0x01, 0x02, 0x10, 0x00, 0x03, 0x0F], np.uint8)
    st = alphabet[rng.integers(0, 6, (nnode, nbin * 4))].view(np.uint32).reshape(nnode, nbin)
    nmat = nmat or nsample + 3
    mat = alphabet[rng.integers(0, 6, (nmat, nbin * 4))].view(np.uint32).reshape(nmat, nbin)
    idx = rng.choice(nmat, nsample, replace=False).astype(np.uint32)
    return st, mat, idx


@pytest.mark.parametrize("nnode", [1, 7, 8, 9, 63, 255, 256, 257, 300, 777, 2100])
def test_vote_root_matches_closed_form(orc, nnode):
    rng = np.random.default_rng(nnode)
    for nbin in (1, 2, 5):
        st, mat, idx = tie_rich(rng, nnode, nbin, 40)
        edge = random_tree_edges(nnode, rng)
        root, nsh = orc.vote_root(st, edge, mat, idx)
        want_root, want_n = np_vote_root2(np_counts(st, edge, mat[idx]))
        assert np.array_equal(nsh, want_n), (nnode, nbin)
        assert np.array_equal(root, want_root), (nnode, nbin)


def test_vote_root_known_answers(orc):
    # 10 nodes in 2 blocks; the sample shares one position with nodes 1, 2
    # and 9 only: in block 0 slot 2 takes slot 0 (stride 2) and then keeps the
    # tie with slot 1 (stride 1), so node 2 stands for block 0 (lane 0) and
    # beats block 1 (lane 1); with every count 0, node 0 wins
    st = np.zeros((10, 1), np.uint32)
    st[1], st[2], st[9] = 0x01, 0x01, 0x01
    edge = np.array([[1, 2], [1, 3], [1, 10]], np.uint32)  # node 0 is the root, its state 0
    mat = np.array([[0x01], [0x10]], np.uint32)
    root, nsh = orc.vote_root(st, edge, mat, [0, 1])
    assert root.tolist() == [2, 0] and nsh.tolist() == [1, 0]
    # 300 nodes (38 blocks): nodes 8 (block 1) and 256 (block 32 = lane 0)
    # tie -- lane 0 beats lane 1 although node 8 has the smaller id
    # (and within a block of 8, node 4 beats node 3: slots 0-3 meet 4-7 first,
    # so slot 0 holds node 4 when it meets slot 1's node 3 or 5)
    st = np.zeros((300, 1), np.uint32)
    st[8] = st[256] = 0x04
    edge = np.array([[1, 9], [1, 257]], np.uint32)
    root, nsh = orc.vote_root(st, edge, np.array([[0x0C]], np.uint32), [0])
    assert root.tolist() == [256] and nsh.tolist() == [1]
    st = np.zeros((8, 1), np.uint32)
    st[3] = st[4] = 0x02
    edge = np.array([[1, 4], [1, 5]], np.uint32)
    root, nsh = orc.vote_root(st, edge, np.array([[0x02]], np.uint32), [0])
    assert root.tolist() == [4] and nsh.tolist() == [1]
    # the edge state is the OR with the parent: node 3's own row is empty
    st = np.zeros((4, 1), np.uint32)
    st[0] = 0x08
    root, nsh = orc.vote_root(st, np.array([[1, 4]], np.uint32), np.array([[0x08]], np.uint32), [0])
    assert root.tolist() == [3] and nsh.tolist() == [1]


def test_find_root_matches_first_edge_max(orc):
    rng = np.random.default_rng(11)
    for nnode in (2, 9, 300, 1000):
        st, mat, idx = tie_rich(rng, nnode, 2, 30)
        edge = random_tree_edges(nnode, rng)
        counts = np_counts(st, edge, mat[idx])
        kids = edge[:, 1].astype(np.int64) - 1
        c = counts[:, kids]
        root, nsh = orc.find_root(st, edge, mat, idx)
        assert np.array_equal(nsh, c.max(axis=1))
        assert np.array_equal(root, kids[np.argmax(c, axis=1)] + 1)
    # no edge: the reference's starting values
    root, nsh = orc.find_root(np.zeros((3, 1), np.uint32), np.zeros((0, 2), np.uint32),
                              np.ones((2, 1), np.uint32), [1])
    assert root.tolist() == [0] and nsh.tolist() == [-1]


# ---------------------------------------------------------------- fasta2bin
IUPAC = {"A": "A", "T": "T", "U": "T", "C": "C", "G": "G", "W": "AT", "M": "AC", "R": "AG", "Y": "CT",
         "K": "GT", "S": "CG", "H": "ACT", "D": "AGT", "V": "ACG", "B": "CGT", "N": "ACGT"}
BIT = {"A": 1, "T": 2, "C": 4, "G": 8}


def np_code_table():
    t = np.full(256, 0x10, np.uint32)
    for ch, bases in IUPAC.items():
        code = sum(BIT[b] for b in bases)
        t[ord(ch)] = t[ord(ch.lower())] = code
    return t


def np_fasta(seqs):
    """fasta2bin by final counts: a column is kept when the OR of its codes'
    base bits has a bit its AND lacks."""
    codes = np_code_table()[seqs]
    base = codes & 0x0F
    orv = np.bitwise_or.reduce(base, axis=0)
    andv = np.bitwise_and.reduce(base, axis=0)
    var = (orv & ~andv & 0x0F) != 0
    valid = (~np.isin(seqs, np.frombuffer(b"-.Nn", np.uint8))).sum(axis=1).astype(np.uint32)
    kept = codes[:, var]
    nvar = kept.shape[1]
    width = (nvar + 3) // 4
    padded = np.zeros((seqs.shape[0], width * 4), np.uint32)
    padded[:, :nvar] = kept
    out = (padded.reshape(seqs.shape[0], width, 4) << np.array([0, 8, 16, 24], np.uint32)).sum(-1, dtype=np.uint64)
    return var, valid, out.astype(np.uint32)


def random_alignment(rng, nseq, length, alphabet):
    a = np.frombuffer(alphabet, np.uint8)
    return a[rng.integers(0, a.size, (nseq, length))]


def test_fasta_known_answer(orc):
    seqs = np.frombuffer(b"ACGTACGANCGT", np.uint8).reshape(3, 4)
    var, valid, out = orc.fasta_encode(seqs)
    # column 0 (A, A, N): T, C, G are in N only; column 3 (T, A, T) differs
    assert var.tolist() == [True, False, False, True]
    assert valid.tolist() == [4, 4, 3]
    assert out.tolist() == [[0x0201], [0x0101], [0x020F]]


@pytest.mark.parametrize("nseq,length", [(1, 9), (2, 1), (3, 17), (40, 64), (200, 333)])
def test_fasta_matches_final_count_rule(orc, nseq, length):
    rng = np.random.default_rng(nseq * 1000 + length)
    for alphabet in (b"ACGT", b"ACGTN-", b"ACGTUWSMKRYBDHVNacgtuwsmkrybdhvn-.X*?0\x00\xff", b"AaN", b"-.X"):
        seqs = random_alignment(rng, nseq, length, alphabet)
        seqs[:, ::7] = seqs[0, ::7]  # constant columns too
        var, valid, out = orc.fasta_encode(seqs)
        wv, wvalid, wout = np_fasta(seqs)
        assert np.array_equal(var, wv), alphabet
        assert np.array_equal(valid, wvalid), alphabet
        assert np.array_equal(out, wout), alphabet


def read_golden():
    from oracle import oracle
    return oracle.read_fasta(open(GOLDEN_FASTA, "rb").read())


def test_fasta_reference_example(orc):
    names, seqs = read_golden()
    assert len(names) == 48 and all(len(s) == 1685 for s in seqs)
    a = np.frombuffer(b"".join(seqs), np.uint8).reshape(len(seqs), -1)
    var, valid, out = orc.fasta_encode(a)
    wv, wvalid, wout = np_fasta(a)
    assert np.array_equal(var, wv) and np.array_equal(valid, wvalid) and np.array_equal(out, wout)
    assert 0 < var.sum() < a.shape[1]
    # RNA: U is encoded as T
    assert (a == ord("U")).any() and not (a == ord("T")).any()


def test_read_fasta_known_answers():
    from oracle import oracle
    text = (b">s1 desc\nAC GT\n\n>  s2 x\nACGA\r\n>s3\n>s4\tz\nNCGT\n"
            b">s5\r\nAAAA")
    names, seqs = oracle.read_fasta(text)
    assert names == [b"s1", b"s2 x", b"s4", b"s5\r"]
    assert seqs == [b"ACGT", b"ACGA", b"NCGT", b"AAAA"]
    # text before the first header is a sequence with an empty name
    names, seqs = oracle.read_fasta(b"AC\nGT\n>b\nTTTT\n")
    assert names == [b"", b"b"] and seqs == [b"ACGT", b"TTTT"]
    with pytest.raises(ValueError, match="same length"):
        oracle.read_fasta(b">a\nAC\n>b\nACG\n")
    with pytest.raises(ValueError, match="No valid sequences"):
        oracle.read_fasta(b">a\n>b\n\n")
