"""The DNA format edge on the GPU: root votes (voteRoot2, findRoot) and
fasta2bin's encoding against the oracle's restatements, through the C ABI and
the drop-in executables, plus size-independent checks at full size."""
import os
import subprocess

import numpy as np
import pytest

from casphy_amd import formats, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "casphy_amd", "bin")
GOLDEN_FASTA = os.path.join(ROOT, "tests", "golden", "run1_1024_head.fasta")


def random_tree_edges(nnode, rng):
    perm = rng.permutation(nnode) + 1
    e = np.array([(perm[rng.integers(0, i)], perm[i]) for i in range(1, nnode)], np.uint32).reshape(-1, 2)
    return e[rng.permutation(e.shape[0])]


def tie_rich(rng, nnode, nbin, nsample):
    alphabet = np.array([0x01, 0x02, 0x10, 0x00, 0x03, 0x0F, 0x04, 0x08], np.uint8)
    k = rng.integers(3, 9)  # few codes: many ties
    st = alphabet[rng.integers(0, k, (nnode, nbin * 4))].view(np.uint32).reshape(nnode, nbin)
    mat = alphabet[rng.integers(0, k, (nsample + 5, nbin * 4))].view(np.uint32).reshape(nsample + 5, nbin)
    idx = rng.choice(nsample + 5, nsample, replace=False).astype(np.uint32)
    return st, mat, idx


@pytest.mark.parametrize("nnode,nbin,nsample", [(1, 1, 3), (9, 1, 40), (257, 2, 100), (300, 1, 64),
                                                (1100, 3, 257), (2500, 8, 130)])
def test_vote_root_tie_rich(hip, orc, nnode, nbin, nsample):
    rng = np.random.default_rng(nnode + nbin)
    st, mat, idx = tie_rich(rng, nnode, nbin, nsample)
    edge = random_tree_edges(nnode, rng)
    root, nsh = hip.vote_root(st, edge, mat, idx)
    want_root, want_n = orc.vote_root(st, edge, mat, idx)
    assert np.array_equal(nsh, want_n)
    assert np.array_equal(root, want_root)
    froot, fn = hip.find_root(st, edge, mat, idx)
    wfroot, wfn = orc.find_root(st, edge, mat, idx)
    assert np.array_equal(fn, wfn) and np.array_equal(froot, wfroot)


@pytest.mark.parametrize("ntip,nsample,ncol", [(60, 40, 48), (700, 300, 250), (3000, 500, 1000)])
def test_vote_root_dna_trees(hip, orc, ntip, nsample, ncol):
    st, edge, mat, idx = synth.vote_inputs(ntip, ntip, nsample, ncol)
    root, nsh = hip.vote_root(st, edge, mat, idx)
    want_root, want_n = orc.vote_root(st, edge, mat, idx)
    assert np.array_equal(nsh, want_n) and np.array_equal(root, want_root)
    froot, fn = hip.find_root(st, edge, mat, idx)
    wfroot, wfn = orc.find_root(st, edge, mat, idx)
    assert np.array_equal(fn, wfn) and np.array_equal(froot, wfroot)


def test_vote_root_edge_cases(hip, orc):
    rng = np.random.default_rng(5)
    st = rng.integers(0, 2**32, (50, 3), dtype=np.uint32) & np.uint32(0x1F1F1F1F)
    edge = random_tree_edges(50, rng)
    mat = rng.integers(0, 2**32, (20, 3), dtype=np.uint32) & np.uint32(0x1F1F1F1F)
    # no sample; samples without a base (every count 0: node 0); repeated rows
    assert hip.vote_root(st, edge, mat, []) [0].size == 0
    gaps = np.full((2, 3), 0x10101010, np.uint32)
    root, nsh = hip.vote_root(st, edge, gaps, [0, 1])
    assert root.tolist() == [0, 0] and nsh.tolist() == [0, 0]
    idx = np.array([3, 3, 7, 3], np.uint32)
    for fn in ("vote_root", "find_root"):
        got = getattr(hip, fn)(st, edge, mat, idx)
        want = getattr(orc, fn)(st, edge, mat, idx)
        assert all(np.array_equal(a, b) for a, b in zip(got, want)), fn
    # no edge: every edge state is 0; findRoot has no candidate
    e0 = np.zeros((0, 2), np.uint32)
    assert [x.tolist() for x in hip.vote_root(st, e0, mat, [1, 2])] == [[0, 0], [0, 0]]
    assert [x.tolist() for x in hip.find_root(st, e0, mat, [1, 2])] == [[0, 0], [-1, -1]]
    # zero-width rows
    z = np.zeros((4, 0), np.uint32)
    assert [x.tolist() for x in hip.vote_root(z, edge[:0], np.zeros((2, 0), np.uint32), [0, 1])] == [[0, 0], [0, 0]]


def test_find_root_r_layout(hip, orc):
    """findRootAVX2Big's R shapes: nodeState (nPositions, nNodes), mat
    (nPositions, nmat) column-major int32, 0-based outIdx."""
    st, edge, mat, idx = synth.vote_inputs(7, 200, 50, 64)
    out = hip.findRootAVX2Big(edge.astype(np.int32), st.T.astype(np.int32), mat.T.astype(np.int32), idx)
    root, nsh = orc.find_root(st, edge, mat, idx)
    assert out.shape == (50, 2) and np.array_equal(out[:, 0], root) and np.array_equal(out[:, 1], nsh)


def test_vote_root_tool(hip, orc, tmp_path):
    st, edge, mat, idx = synth.vote_inputs(9, 400, 120, 200)
    pfx = str(tmp_path / "x__M1")
    matfile = str(tmp_path / "x.bin")
    formats.write_bin(matfile, mat)
    formats.write_bin(pfx + ".nodeState", st)
    formats.write_bin(pfx + ".edge", edge)
    formats.write_bin(pfx + ".outmatIdx", idx.reshape(1, -1))
    r = subprocess.run([os.path.join(BIN, "voteRoot2"), matfile, pfx], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    got = formats.read_bin(pfx + ".rootVote")
    root, nsh = orc.vote_root(st, edge, mat, idx)
    assert got.shape == (2, 120)
    assert np.array_equal(got[0], root) and np.array_equal(got[1], nsh)


# ---------------------------------------------------------------- fasta2bin
@pytest.mark.parametrize("nseq,length", [(1, 5), (2, 16), (3, 33), (64, 1000), (513, 70), (1500, 2049)])
def test_fasta_encode(hip, orc, nseq, length):
    rng = np.random.default_rng(nseq + length)
    for alphabet in (b"ACGT-", b"ACGTUWSMKRYBDHVNacgtuwsmkrybdhvn-.X*?0\x00\xff", b"AaNn"):
        a = np.frombuffer(alphabet, np.uint8)
        seqs = a[rng.integers(0, a.size, (nseq, length))]
        seqs[:, ::5] = seqs[0, ::5]
        got = hip.fasta_encode(seqs)
        want = orc.fasta_encode(seqs)
        for g, w in zip(got, want):
            assert np.array_equal(g, w), alphabet


def test_fasta_encode_staged_copies(hip, orc):
    """67,109 x 1,001 characters: the upload (67.2 MB) and the read-back of
    the codes (67.1 MB) are above 64 MiB and go through the pinned staging
    buffers in 32 MiB chunks with a partial last chunk; identical to the
    oracle's encoding."""
    rng = np.random.default_rng(91)
    a = np.frombuffer(b"ACGT-N", np.uint8)
    seqs = a[rng.integers(0, a.size, (67109, 1001))]
    assert seqs.nbytes >= 64 << 20
    got = hip.fasta_encode(seqs)
    assert got[2].nbytes >= 64 << 20
    want = orc.fasta_encode(seqs)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_fasta_encode_full_size(hip):
    """100,000 x 1,000 characters of the JC69 alignment: each kept column's
    codes decode back to the characters, dropped columns are constant, valid
    counts add up -- checked column-wise in numpy."""
    codes = synth.dna_matrix(seed=4, n=100000, ncol=1000)
    b = codes.view(np.uint8).reshape(codes.shape[0], -1)[:, :1000]
    chars = np.frombuffer(b"?ATTCCCCGGGGGGGG-", np.uint8)  # code -> a character of that code
    seqs = chars[np.where(b == 0x10, 16, b)]
    seqs[:, 10] = ord("N")  # a constant N column is dropped
    var, valid, out = hip.fasta_encode(seqs)
    base = np.where(b == 0x10, 0, b)
    base[:, 10] = 0x0F
    want_var = (np.bitwise_or.reduce(base, 0) & ~np.bitwise_and.reduce(base, 0)) != 0
    assert np.array_equal(var, want_var) and not var[10]
    assert np.array_equal(valid, (seqs != ord("-")).sum(1) - 1)
    dec = out.view(np.uint8).reshape(out.shape[0], -1)[:, : int(var.sum())]
    assert np.array_equal(dec, np.where(b == 0x10, 0x10, base)[:, var])


def fasta_text(names, seqs, wrap=None):
    parts = []
    for n, s in zip(names, seqs):
        parts.append(b">" + n + b"\n")
        if wrap:
            parts += [s[i:i + wrap] + b"\n" for i in range(0, len(s), wrap)]
        else:
            parts.append(s + b"\n")
    return b"".join(parts)


def check_fasta2bin(orc, tmp_path, text):
    from oracle import oracle
    src = tmp_path / "in.fa"
    src.write_bytes(text)
    pfx = str(tmp_path / "out")
    r = subprocess.run([os.path.join(BIN, "fasta2bin"), str(src), pfx], capture_output=True, timeout=600)
    assert r.returncode == 0, r.stderr
    names, seqs = oracle.read_fasta(text)
    a = np.frombuffer(b"".join(seqs), np.uint8).reshape(len(seqs), -1)
    var, valid, out = orc.fasta_encode(a)
    raw = np.fromfile(pfx + ".bin", np.uint32)
    assert raw[0] == len(seqs) and raw[1] == out.shape[1]
    assert np.array_equal(raw[2:].reshape(out.shape), out)
    want_names = b"SequenceName,ValidBaseCount\n" + b"".join(n + b"," + str(v).encode() + b"\n"
                                                            for n, v in zip(names, valid))
    assert open(pfx + "_names.txt", "rb").read() == want_names
    want_pos = b"Position,IsVariable\n" + b"".join(str(i).encode() + (b",Yes\n" if v else b",No\n")
                                                  for i, v in enumerate(var))
    assert open(pfx + "_positions.txt", "rb").read() == want_pos


def test_fasta2bin_tool_reference_example(orc, tmp_path):
    check_fasta2bin(orc, tmp_path, open(GOLDEN_FASTA, "rb").read())


def test_fasta2bin_tool_text_forms(orc, tmp_path):
    rng = np.random.default_rng(2)
    a = np.frombuffer(b"ACGTNacgtn-.RYU", np.uint8)
    seqs = [bytes(a[rng.integers(0, a.size, 301)]) for _ in range(37)]
    names = [b"seq%d extra words" % i for i in range(37)]
    names[3] = b"  leading blank"
    names[4] = b"tab\tcut"
    text = fasta_text(names, seqs, wrap=60)
    text = text.replace(b"\n", b"\r\n", 40)  # CR before some line ends
    text = b"\n\n" + text.replace(b">seq7", b">ghost\n>seq7")  # empty lines, a record without sequence
    check_fasta2bin(orc, tmp_path, text)
    check_fasta2bin(orc, tmp_path, b">only\nACGTN\n")  # one sequence: nothing varies
