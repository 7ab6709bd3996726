"""The oracle's fasta2bin restatement (oracle.read_fasta + orc_fasta_encode)
against the reference's own fasta2bin (cu/fasta2bin.cu), compiled from its
source into oracle/_ref/ by `make -C oracle ref`: every output file, byte for
byte, on the reference's example alignment, random IUPAC alignments and the
FASTA text forms the reader handles (wrapped lines, CR line ends, blank lines,
names with blanks, a record with no sequence).  Skipped where the reference
binary was not built (its source is not on every host)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "fasta2bin")
GOLDEN_FASTA = os.path.join(ROOT, "tests", "golden", "run1_1024_head.fasta")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_BIN), reason="oracle/_ref/fasta2bin not built")


def oracle_files(orc, text: bytes):
    """The three fasta2bin outputs (.bin, _names.txt, _positions.txt) made
    from the oracle's reader and encoder."""
    names, seqs = orc.read_fasta(text)
    a = np.frombuffer(b"".join(seqs), np.uint8).reshape(len(seqs), -1)
    var, valid, out = orc.fasta_encode(a)
    binf = np.array([len(seqs), out.shape[1]], np.uint32).tobytes() + np.ascontiguousarray(out).tobytes()
    namesf = b"SequenceName,ValidBaseCount\n" + b"".join(n + b"," + str(int(v)).encode() + b"\n"
                                                         for n, v in zip(names, valid))
    posf = b"Position,IsVariable\n" + b"".join(str(i).encode() + (b",Yes\n" if v else b",No\n")
                                               for i, v in enumerate(var))
    return binf, namesf, posf


def ref_files(tmp_path, text: bytes):
    src = tmp_path / "in.fa"
    src.write_bytes(text)
    pfx = str(tmp_path / "ref")
    r = subprocess.run([REF_BIN, str(src), pfx], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return tuple(open(pfx + suf, "rb").read() for suf in (".bin", "_names.txt", "_positions.txt"))


def fasta_text(names, seqs, wrap=None):
    parts = []
    for n, s in zip(names, seqs):
        parts.append(b">" + n + b"\n")
        if wrap:
            parts += [s[i:i + wrap] + b"\n" for i in range(0, len(s), wrap)]
        else:
            parts.append(s + b"\n")
    return b"".join(parts)


def check(orc, tmp_path, text):
    got, want = ref_files(tmp_path, text), oracle_files(orc, text)
    for g, w, what in zip(got, want, (".bin", "_names.txt", "_positions.txt")):
        assert g == w, what


def test_reference_example(orc, tmp_path):
    check(orc, tmp_path, open(GOLDEN_FASTA, "rb").read())


@pytest.mark.parametrize("nseq,length,alphabet", [(2, 1, b"AC"), (3, 17, b"ACGTN-"), (40, 64, b"ACGT"),
                                                  (37, 301, b"ACGTNacgtn-.RYUWSMKBDHV"),
                                                  (64, 200, b"ACGTUWSMKRYBDHVNacgtuwsmkrybdhvn-.X*?0"),
                                                  (9, 50, b"-.X"), (5, 33, b"AaN")])
def test_random_alignments(orc, tmp_path, nseq, length, alphabet):
    rng = np.random.default_rng(nseq * 7 + length)
    a = np.frombuffer(alphabet, np.uint8)
    seqs = [bytes(a[rng.integers(0, a.size, length)]) for _ in range(nseq)]
    seqs[0] = seqs[0][:length // 2] + seqs[-1][length // 2:]  # shared halves: constant columns too
    check(orc, tmp_path, fasta_text([b"s%d" % i for i in range(nseq)], seqs, wrap=60))


def test_text_forms(orc, tmp_path):
    rng = np.random.default_rng(2)
    a = np.frombuffer(b"ACGTNacgtn-.RYU", np.uint8)
    seqs = [bytes(a[rng.integers(0, a.size, 301)]) for _ in range(37)]
    names = [b"seq%d extra words" % i for i in range(37)]
    names[3] = b"  leading blank"
    names[4] = b"tab\tcut"
    text = fasta_text(names, seqs, wrap=60)
    text = text.replace(b"\n", b"\r\n", 40)  # CR before some line ends
    text = b"\n\n" + text.replace(b">seq7", b">ghost\n>seq7")  # empty lines, a record without sequence
    check(orc, tmp_path, text)
    check(orc, tmp_path, b">only\nACGTN\n")  # one sequence: nothing varies
    check(orc, tmp_path, b">a\nACGT\n>b\nACGA\n")  # no final blank line handling needed
    check(orc, tmp_path, b">a\nAC\nGT\n\n>b\nAC\n\nGA")  # blank line inside a record, no final newline
