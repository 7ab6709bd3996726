"""Builds the golden fixtures in tests/golden/ from the reference's own example
inputs (read as data, never executed):

  data/mat1k.rda        -> indel_mat1k.npz   (1,000 x 30 Cas9 indel matrix)
  data/mat_binary.rda   -> binary_mat.npz    (500 x 3,000 base-editor matrix)
  examples/run1_1024.fasta -> dna_run1.npz   (first 192 sequences, fasta2bin codes)

Each fixture holds the input in the layout the executables read plus the
oracle's outputs for it (join edges, cladewise order, radius list lengths and
ids, tryRadius counts).  Run once in a checkout that has the reference next to
it:  python tests/golden/make_golden.py /root/reference
The .rda files are gzip'd R XDR serialisations; the small reader below handles
the object types they use (pairlists, symbols, strings, integer / logical /
double vectors with attributes).
"""
from __future__ import annotations

import gzip
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


# ------------------------------------------------------------------ R XDR
class XDR:
    def __init__(self, data: bytes):
        self.d = data
        self.p = 0
        self.refs = []

    def int(self):
        v = struct.unpack_from(">i", self.d, self.p)[0]
        self.p += 4
        return v

    def ints(self, n):
        v = np.frombuffer(self.d, ">i4", n, self.p).astype(np.int64)
        self.p += 4 * n
        return v

    def doubles(self, n):
        v = np.frombuffer(self.d, ">f8", n, self.p)
        self.p += 8 * n
        return v

    def item(self):
        flags = self.int()
        t = flags & 0xFF
        has_attr, has_tag = bool(flags & (1 << 9)), bool(flags & (1 << 10))
        if t == 254 or t == 0:              # NILVALUE_SXP / NILSXP
            return None
        if t == 255:                        # REFSXP
            idx = flags >> 8
            if idx == 0:
                idx = self.int()
            return self.refs[idx - 1]
        if t == 1:                          # SYMSXP
            name = self.item()
            self.refs.append(("sym", name))
            return ("sym", name)
        if t == 9:                          # CHARSXP
            n = self.int()
            if n == -1:
                return None
            s = self.d[self.p:self.p + n].decode("latin-1")
            self.p += n
            return s
        if t == 2:                          # LISTSXP (pairlist)
            out = []
            while True:
                attr = self.item() if has_attr else None
                tag = self.item() if has_tag else None
                car = self.item()
                out.append((tag[1] if tag else None, car))
                nxt = self.int()
                t = nxt & 0xFF
                if t != 2:
                    self.p -= 4
                    self.item()
                    break
                has_attr, has_tag = bool(nxt & (1 << 9)), bool(nxt & (1 << 10))
            return out
        if t in (10, 13):                   # LGLSXP, INTSXP
            v = self.ints(self.int())
        elif t == 14:                       # REALSXP
            v = self.doubles(self.int())
        elif t == 16 or t == 19:            # STRSXP, VECSXP
            v = [self.item() for _ in range(self.int())]
        else:
            raise ValueError(f"unsupported SEXP type {t}")
        attrs = dict((k, a) for k, a in self.item()) if has_attr else {}
        return ("vec", v, attrs)


def read_rda(path):
    raw = gzip.open(path).read()
    assert raw[:5] == b"RDX3\n" and raw[5:7] == b"X\n", "not an XDR RDX3 file"
    x = XDR(raw[7:])
    version = x.int()
    x.int(); x.int()
    if version == 3:
        nlen = x.int()                      # native encoding name
        x.p += nlen
    return dict(x.item())


def r_matrix(obj):
    _, v, attrs = obj
    nr, nc = (int(d) for d in attrs["dim"][1])
    return np.asarray(v).reshape(nc, nr).T  # column-major -> (nrow, ncol)


# ------------------------------------------------------------------ fasta2bin
CODES = {c: 0x10 for c in range(256)}
for s, v in (("A", 1), ("T", 2), ("U", 2), ("C", 4), ("G", 8), ("W", 3), ("M", 5), ("R", 9), ("Y", 6), ("K", 10),
             ("S", 12), ("H", 7), ("D", 11), ("V", 13), ("B", 14), ("N", 15)):
    CODES[ord(s)] = v
    CODES[ord(s.lower())] = v
BASES = {1: [0], 2: [1], 4: [2], 8: [3], 3: [0, 1], 5: [0, 2], 9: [0, 3], 6: [1, 2], 10: [1, 3], 12: [2, 3],
         7: [0, 1, 2], 11: [0, 1, 3], 13: [0, 2, 3], 14: [1, 2, 3], 15: [0, 1, 2, 3]}


def read_fasta(path):
    seqs, cur = [], []
    for line in open(path):
        line = line.strip()
        if not line:
            continue
        if line[0] == ">":
            if cur:
                seqs.append("".join(cur))
            cur = []
        else:
            cur.append("".join(line.split()))
    if cur:
        seqs.append("".join(cur))
    return seqs


def fasta2bin_codes(seqs):
    """Byte codes after the variable-position filter of cu/fasta2bin.cu:47-93:
    a position is kept once some nucleotide count is strictly between 0 and
    the number of sequences checked so far (IUPAC codes count every base,
    anything else counts as 'other')."""
    codes = np.array([[CODES[ord(c)] for c in s] for s in seqs], dtype=np.uint8)
    keep = np.zeros(codes.shape[1], bool)
    for p in range(codes.shape[1]):
        cnt = [0, 0, 0, 0]
        for k in range(codes.shape[0]):
            for b in BASES.get(int(codes[k, p]), []):
                cnt[b] += 1
            if any(0 < c < k + 1 for c in cnt):
                keep[p] = True
                break
    return codes[:, keep]


def main(ref):
    from casphy_amd import synth
    from oracle import oracle

    # indel: mat1k (cells x sites), regularised as shared.mutation.join does
    m = read_rda(os.path.join(ref, "data", "mat1k.rda"))
    mat = synth.regularize(r_matrix(m["mat1k"]).astype(np.int32))
    w = np.ones((max(1, int(mat.max())), mat.shape[1]), np.float32)
    e, ns, br = oracle.join_indel(mat, w)
    eo, order = oracle.cladewise(e, mat.shape[0])
    np.savez_compressed(os.path.join(HERE, "indel_mat1k.npz"), mat=mat, edge=e, node_state=ns, branch=br,
                        edge2=eo, order=order)

    # binary: mat_binary (cells x sites 0/1) -> to.bit (one bit per site)
    m = read_rda(os.path.join(ref, "data", "mat_binary.rda"))
    b = r_matrix(m["mat_binary"]) > 0
    packed = synth.pack_bits(b)
    e, ns, nb = oracle.join_bit(packed, ancestral_type=1)
    eo, order = oracle.cladewise(e, packed.shape[0])
    radii = np.array([1, 5, 20], np.int32)
    lists = {f"counts_r{r}": oracle.radius_bit(packed, int(r))[0] for r in radii}
    ids = {f"idx_r{r}": oracle.radius_bit(packed, int(r))[1] for r in radii if r >= 5}
    np.savez_compressed(os.path.join(HERE, "binary_mat.npz"), mat=packed, edge=e, node_state=ns, edge2=eo,
                        order=order, radii=radii, **lists, **ids)

    # DNA: run1_1024.fasta, first 192 sequences through the fasta2bin filter
    seqs = read_fasta(os.path.join(ref, "examples", "run1_1024.fasta"))[:192]
    codes = fasta2bin_codes(seqs)
    tb = synth.pack_bytes4(codes)
    q = np.arange(0, 192, 7, dtype=np.uint32)
    radii = np.array([200, int(codes.shape[1] * 0.6), int(codes.shape[1] * 0.8)], np.int32)
    out = {"mat": tb, "radii": radii, "q": q}
    for r in radii:
        c, i = oracle.radius_dna(tb, int(r))
        out[f"counts_r{r}"] = c
        out[f"idx_r{r}"] = i
        out[f"try_r{r}"] = oracle.try_radius_dna(tb, int(r), q)
    np.savez_compressed(os.path.join(HERE, "dna_run1.npz"), **out)
    print("fixtures written:", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
