"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
The product (casphy_amd) never imports it.

Array conventions follow the reference's on-disk layouts: binary and DNA
matrices are cell-major (n, W) uint32, indel matrices (n, ncol) int32 with
-1 missing / 0 unmutated / k>0 allele, weights float32 (the executables convert
the .w doubles to float before summing, cu/radiusNeighbor.cu:240-243).

Three restatements are pure Python here rather than C: read_fasta (the FASTA
line rules of cu/fasta2bin.cu:121-187), read_newick (ape's read.tree
numbering) and tbe_support (the TBE formula).  Their pins are in
tests/test_oracle_dna_edge.py and tests/test_oracle_trees.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        u32, f32, i32, u64 = C.c_uint32, C.c_float, C.c_int, C.c_uint64
        vp = C.c_void_p
        L.orc_score_bit.restype = f32
        L.orc_score_bit.argtypes = [u32p, u32p, u32, f32p]
        L.orc_score_bit_matrix.argtypes = [u32p, u32, u32, f32p, f32p]
        L.orc_score_indel.restype = f32
        L.orc_score_indel.argtypes = [i32p, i32p, u32, f32p]
        L.orc_score_dna.restype = u32
        L.orc_score_dna.argtypes = [u32p, u32p, u32]
        L.orc_radius_bit_rows.restype = i32
        L.orc_radius_bit_rows.argtypes = [u32p, u32, u32, f32p, f32, u32, u32, u32p, vp, u64,
                                          C.POINTER(C.c_uint64)]
        L.orc_radius_dna_rows.restype = i32
        L.orc_radius_dna_rows.argtypes = [u32p, u32, u32, u32, u32, u32, u32p, vp, u64,
                                          C.POINTER(C.c_uint64)]
        L.orc_try_radius_bit.argtypes = [u32p, u32, u32, f32p, f32, u32p, u32, u32p]
        L.orc_try_radius_dna.argtypes = [u32p, u32, u32, u32, u32p, u32, u32p]
        L.orc_score_indel_matrix.argtypes = [i32p, u32, u32, f32p, f32p]
        L.orc_join_bit.restype = i32
        L.orc_join_bit.argtypes = [u32p, u32, u32, f32p, i32, u32p, u32p, u32p]
        L.orc_join_bit_lazy.restype = i32
        L.orc_join_bit_lazy.argtypes = [u32p, u32, u32, f32p, i32, u32p, u32p, u32p]
        L.orc_join_indel.restype = i32
        L.orc_join_indel.argtypes = [i32p, u32, u32, f32p, u32p, u32, f32, u32p, i32p, u32p]
        L.orc_join_indel_lazy.restype = i32
        L.orc_join_indel_lazy.argtypes = [i32p, u32, u32, f32p, u32p, u32, f32, u32p, i32p, u32p]
        L.orc_cladewise.argtypes = [u32p, u32, u32p, u32p]
        L.orc_transfer.argtypes = [u32p, u32, u32p, u32, u32, u32, u32p, u32p, u32p, u32p]
        L.orc_list_shared.restype = u32
        L.orc_list_shared.argtypes = [u32p, u32, u32p, u32]
        L.orc_tip_merge.restype = i32
        L.orc_tip_merge.argtypes = [u32p, C.POINTER(C.c_uint64), u32p, u32, C.c_double, i32p]
        L.orc_node_keep.restype = u32
        L.orc_clade_bitsets.restype = u32
        L.orc_clade_bitsets.argtypes = [i32p, i32p, u32, u32, u32, u32p, i32p]
        L.orc_node_keep.argtypes = [i32, u32p, u32, u32p, u32, u32, u32p, u32, u32, u32p]
        L.orc_vote_root.argtypes = [u32p, u32, u32, u32p, u32, u32p, u32, u32p, u32p]
        L.orc_find_root.argtypes = [u32p, u32, u32, u32p, u32, u32p, u32, i32p, i32p]
        u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
        L.orc_fasta_encode.restype = u64
        L.orc_fasta_encode.argtypes = [u8p, u32, u64, u8p, u32p, u32p]
        _lib = L
    return _lib


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def unit_weights_bit(W: int) -> np.ndarray:
    return np.ones(32 * W, dtype=np.float32)


# ---------------------------------------------------------------- scoring
def score_bit(q, t, w=None) -> float:
    q, t = _u32(q), _u32(t)
    w = unit_weights_bit(q.size) if w is None else _f32(w)
    return float(lib().orc_score_bit(q, t, q.size, w))


def score_bit_matrix(mat, w=None) -> np.ndarray:
    mat = _u32(mat)
    n, W = mat.shape
    w = unit_weights_bit(W) if w is None else _f32(w)
    S = np.empty((n, n), dtype=np.float32)
    lib().orc_score_bit_matrix(mat, n, W, w, S)
    return S


def score_indel(a, b, w) -> float:
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    return float(lib().orc_score_indel(a, b, a.size, _f32(w)))


def score_indel_matrix(mat, w) -> np.ndarray:
    """All pairs of score_indel, the diagonal included (float32)."""
    mat = np.ascontiguousarray(mat, dtype=np.int32)
    n, ncol = mat.shape
    S = np.empty((n, n), dtype=np.float32)
    lib().orc_score_indel_matrix(mat, n, ncol, _f32(w), S)
    return S


def score_dna(a, b) -> int:
    a, b = _u32(a), _u32(b)
    return int(lib().orc_score_dna(a, b, a.size))


# ---------------------------------------------------------------- radius
def _radius_rows(kind, mat, radius, w, rows=None):
    mat = _u32(mat)
    n, W = mat.shape
    i0, i1 = (0, n) if rows is None else rows
    counts = np.zeros(i1 - i0, dtype=np.uint32)
    pos = C.c_uint64(0)
    L = lib()
    if kind == "bit":
        w = unit_weights_bit(W) if w is None else _f32(w)
        L.orc_radius_bit_rows(mat, n, W, w, float(radius), i0, i1, counts, None, 0, C.byref(pos))
    else:
        L.orc_radius_dna_rows(mat, n, W, int(radius), i0, i1, counts, None, 0, C.byref(pos))
    total = int(counts.sum(dtype=np.uint64))
    idx = np.zeros(max(total, 1), dtype=np.uint32)
    pos = C.c_uint64(0)
    if kind == "bit":
        rc = L.orc_radius_bit_rows(mat, n, W, w, float(radius), i0, i1, counts,
                                   idx.ctypes.data_as(C.c_void_p), total, C.byref(pos))
    else:
        rc = L.orc_radius_dna_rows(mat, n, W, int(radius), i0, i1, counts,
                                   idx.ctypes.data_as(C.c_void_p), total, C.byref(pos))
    assert rc == 0
    return counts, idx[:total]


def radius_pass(kind, mat, radius, rows, w=None):
    """One pass of the reference's per-query radius loop over query rows
    rows = (i0, i1): (counts, idx) with the ids written as they are found
    (radiusWeightedParallel / radiusWeightedDNAParallel, src/functions.cpp:
    1825-1857 / 4021-4036, visit every target once per query).  Used to time
    the CPU path; radius_bit / radius_dna size their output first."""
    mat = _u32(mat)
    n, W = mat.shape
    i0, i1 = rows
    counts = np.zeros(i1 - i0, dtype=np.uint32)
    idx = np.empty((i1 - i0) * n, dtype=np.uint32)
    pos = C.c_uint64(0)
    L = lib()
    if kind == "bit":
        w = unit_weights_bit(W) if w is None else _f32(w)
        rc = L.orc_radius_bit_rows(mat, n, W, w, float(radius), i0, i1, counts, idx.ctypes.data_as(C.c_void_p),
                                   idx.size, C.byref(pos))
    else:
        rc = L.orc_radius_dna_rows(mat, n, W, int(radius), i0, i1, counts, idx.ctypes.data_as(C.c_void_p),
                                   idx.size, C.byref(pos))
    assert rc == 0
    return counts, idx[:pos.value]


def radius_bit(mat, radius, w=None, rows=None):
    """(counts, idx) = the .nNeighbor counts and the concatenated .idx lists."""
    return _radius_rows("bit", mat, radius, w, rows)


def radius_dna(mat, radius, rows=None):
    return _radius_rows("dna", mat, radius, None, rows)


def try_radius_bit(mat, radius, q, w=None) -> np.ndarray:
    mat, q = _u32(mat), _u32(q)
    n, W = mat.shape
    w = unit_weights_bit(W) if w is None else _f32(w)
    out = np.zeros(q.size, dtype=np.uint32)
    lib().orc_try_radius_bit(mat, n, W, w, float(radius), q, q.size, out)
    return out


def try_radius_dna(mat, radius, q) -> np.ndarray:
    mat, q = _u32(mat), _u32(q)
    n, W = mat.shape
    out = np.zeros(q.size, dtype=np.uint32)
    lib().orc_try_radius_dna(mat, n, W, int(radius), q, q.size, out)
    return out


# ---------------------------------------------------------------- joins
def join_bit(mat, w=None, ancestral_type=1, lazy=False):
    """Returns (edge (2n-2,2) raw join order, node_state (2n-1,W), node_bistate)."""
    mat = _u32(mat)
    n, W = mat.shape
    w = unit_weights_bit(W) if w is None else _f32(w)
    edge = np.zeros((2 * n - 2, 2), dtype=np.uint32)
    ns = np.zeros((2 * n - 1, W), dtype=np.uint32)
    nb = np.zeros((2 * n - 1, W), dtype=np.uint32)
    fn = lib().orc_join_bit_lazy if lazy else lib().orc_join_bit
    rc = fn(mat, n, W, w, int(ancestral_type), edge, ns, nb)
    if rc != 0:
        raise RuntimeError(f"oracle join_bit rc={rc}")
    return edge, ns, nb


def join_indel(mat, w, tip_size=None, min_size=0, min_ratio=2.0, lazy=False):
    """Returns (edge (2n-2,2), node_state (2n-1,ncol) int32, branch (2n-2,))."""
    mat = np.ascontiguousarray(mat, dtype=np.int32)
    n, ncol = mat.shape
    ts = np.ones(n, dtype=np.uint32) if tip_size is None else _u32(tip_size)
    edge = np.zeros((2 * n - 2, 2), dtype=np.uint32)
    ns = np.zeros((2 * n - 1, ncol), dtype=np.int32)
    br = np.zeros(2 * n - 2, dtype=np.uint32)
    fn = lib().orc_join_indel_lazy if lazy else lib().orc_join_indel
    rc = fn(mat, n, ncol, _f32(w), ts, int(min_size), float(min_ratio), edge, ns, br)
    if rc != 0:
        raise RuntimeError(f"oracle join_indel rc={rc}")
    return edge, ns, br


def cladewise(edge_raw, n):
    """(edge_out (2n-2,2) relabelled DFS order, internal_order (n-1,) raw ids)."""
    edge_raw = _u32(edge_raw)
    eo = np.zeros((2 * n - 2, 2), dtype=np.uint32)
    io = np.zeros(max(n - 1, 1), dtype=np.uint32)
    lib().orc_cladewise(edge_raw, n, eo, io)
    return eo, io[: n - 1]


# ---------------------------------------------------------------- transfer
def transfer(Y, X, ntip):
    Y, X = _u32(Y), _u32(X)
    ny, ncol = Y.shape
    nx = X.shape[0]
    ymin = np.zeros(ny, np.uint32)
    ysize = np.zeros(ny, np.uint32)
    xmin = np.zeros(nx, np.uint32)
    xsize = np.zeros(nx, np.uint32)
    lib().orc_transfer(Y, ny, X, nx, ncol, int(ntip), ymin, ysize, xmin, xsize)
    return ymin, ysize, xmin, xsize


# ---------------------------------------------------------------- tipMerge
def list_shared(a, b) -> int:
    a, b = _u32(a), _u32(b)
    return int(lib().orc_list_shared(a, a.size, b, b.size))


def tip_merge(nnb, idx, p: float) -> np.ndarray:
    """Cluster labels (int32 per tip) of cu/tipMerge.cu for the concatenated
    ascending neighbour lists idx with lengths nnb."""
    nnb, idx = _u32(nnb), _u32(idx)
    n = nnb.size
    off = np.zeros(n + 1, dtype=np.uint64)
    off[1:] = np.cumsum(nnb, dtype=np.uint64)
    out = np.zeros(max(n, 1), dtype=np.int32)
    rc = lib().orc_tip_merge(nnb, off.ctypes.data_as(C.POINTER(C.c_uint64)), idx if idx.size else _u32([0]), n,
                             float(p), out)
    if rc != 0:
        raise RuntimeError(f"oracle tip_merge rc={rc}")
    return out[:n]


# ---------------------------------------------------------------- nodeKeep
def node_keep(edge, state, outmat, cutoff: int, indel: bool = False) -> np.ndarray:
    """Kept internal node ids of nodeKeepBit / nodeKeepIndel: edge (nedge, 2)
    0-based, state (nnode, width), outmat (nout, width), all u32."""
    edge, state, outmat = _u32(edge), _u32(state), _u32(outmat)
    nnode, width = state.shape
    nout = outmat.shape[0] if outmat.size else 0
    kept = np.zeros(max(nnode, 1), np.uint32)
    k = lib().orc_node_keep(int(indel), edge.reshape(-1) if edge.size else _u32([0]), edge.shape[0], state, nnode,
                            width, outmat if outmat.size else _u32([0]), nout, int(cutoff), kept)
    return kept[:k]


# ---------------------------------------------------------------- clade bitsets
def clade_bitsets(edge, nnode: int, ntip: int):
    """(rows (nnode-1, ceil(ntip/32)) u32, sizes (nnode-1,) int32) of
    nodeDesBig for R's 1-based edge matrix (nedge, 2)."""
    e = np.ascontiguousarray(edge, dtype=np.int32)
    anc, des = np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1])
    nbin = ((ntip - 1) >> 5) + 1
    rows = np.zeros((max(nnode - 1, 1), nbin), np.uint32)
    sizes = np.zeros(max(nnode - 1, 1), np.int32)
    k = lib().orc_clade_bitsets(anc, des, e.shape[0], nnode, ntip, rows, sizes)
    return rows[:k], sizes[: max(nnode - 1, 0)]


# ---------------------------------------------------------------- DNA format edge
def _root_inputs(state, edge, mat, idx):
    state = _u32(state)
    nnode, nbin = state.shape
    edge = _u32(edge).reshape(-1, 2)
    mat = _u32(mat).reshape(-1, nbin)
    samples = _u32(mat[np.asarray(idx, dtype=np.int64)]) if len(idx) else np.zeros((0, nbin), np.uint32)
    return state, nnode, nbin, edge, samples


def vote_root(state, edge, mat, idx):
    """voteRoot2: (root 0-based, nshared) per sample row mat[idx[s]]; state
    (nnode, nbin) u32 byte codes, edge (nedge, 2) 1-based."""
    state, nnode, nbin, edge, samples = _root_inputs(state, edge, mat, idx)
    ns = samples.shape[0]
    root, nsh = np.zeros(max(ns, 1), np.uint32), np.zeros(max(ns, 1), np.uint32)
    lib().orc_vote_root(state, nnode, nbin, edge.reshape(-1) if edge.size else _u32([0]), edge.shape[0],
                        samples if ns else _u32([0]), ns, root, nsh)
    return root[:ns], nsh[:ns]


def find_root(state, edge, mat, idx):
    """findRoot / findRootAVX2Big: (root 1-based, nshared) int32 per sample."""
    state, nnode, nbin, edge, samples = _root_inputs(state, edge, mat, idx)
    ns = samples.shape[0]
    root, nsh = np.zeros(max(ns, 1), np.int32), np.zeros(max(ns, 1), np.int32)
    lib().orc_find_root(state, nnode, nbin, edge.reshape(-1) if edge.size else _u32([0]), edge.shape[0],
                        samples if ns else _u32([0]), ns, root, nsh)
    return root[:ns], nsh[:ns]


def fasta_encode(seqs):
    """fasta2bin on an (nseq, len) uint8 array of equal-length sequences:
    (is_variable bool (len,), valid (nseq,) u32, codes (nseq, width) u32)."""
    seqs = np.ascontiguousarray(seqs, dtype=np.uint8)
    nseq, length = seqs.shape
    isv = np.zeros(max(length, 1), np.uint8)
    valid = np.zeros(max(nseq, 1), np.uint32)
    out = np.zeros(max(nseq * ((length + 3) // 4), 1), np.uint32)
    width = lib().orc_fasta_encode(seqs if seqs.size else np.zeros(1, np.uint8), nseq, length, isv, valid, out)
    return isv[:length].astype(bool), valid[:nseq], out[: nseq * width].reshape(nseq, width)


def read_fasta(text: bytes):
    """cu/fasta2bin.cu:121-187 (read_fasta): (names, sequences) of a FASTA
    text, line by line: empty lines skipped; a '>' line ends the running
    sequence (kept only when non-empty) and names the next one (the text after
    '>' from its first non-blank character, cut at the first blank when that
    blank comes after it); other lines join the running sequence with their
    whitespace removed.  Raises ValueError as the reference throws."""
    names, seqs = [], []
    name, cur = b"", b""
    lines = text.split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()  # getline: a final newline ends the last line
    ws = b" \t\n\v\f\r"
    for line in lines:
        if not line:
            continue
        if line[:1] == b">":
            if cur:
                seqs.append(cur)
                names.append(name)
                cur = b""
            name = line[1:]
            first = next((i for i, c in enumerate(name) if c not in b" \t"), None)
            blank = next((i for i, c in enumerate(name) if c in b" \t"), None)
            if first is not None:
                name = name[first:] if blank is None or blank < first else name[first:blank]
        else:
            cur += bytes(c for c in line if c not in ws)
    if cur:
        seqs.append(cur)
        names.append(name)
    if not seqs:
        raise ValueError("No valid sequences found in file")
    if any(len(x) != len(seqs[0]) for x in seqs):
        raise ValueError("All sequences must have the same length for variant filtering")
    return names, seqs


# ---------------------------------------------------------------- trees from Newick
def read_newick(text, tip_labels=None):
    """Newick trees numbered as ape's read.tree (ape >= 5.0, a dependency of
    CasPhy, DESCRIPTION:10-12) numbers them: tips 1..n by order of
    appearance (or by position in tip_labels), internal nodes n+1.. by the
    order of their '(' (root first), edges in that preorder.  Pure Python:
    the text is cut into tokens, then an explicit stack of open nodes builds
    the edge list.  Returns [(edge (nedge, 2) int32, nnode, ntip, labels)]."""
    import re
    if isinstance(text, bytes):
        text = text.decode("utf-8", "surrogateescape")
    tok = re.compile(r"\s+|\[[^\]]*\]|'(?:[^']|'')*'|[(),;]|:[^(),;:\[\s]*|[^(),;:\[\s']+")
    trees, toks = [], []
    for m in tok.finditer(text):
        t = m.group(0)
        if t[0].isspace() or t[0] == "[":
            continue
        toks.append(t)
        if t == ";":
            trees.append(toks)
            toks = []
    if any(t not in ";" for t in toks):
        raise ValueError("newick: missing ';'")
    out = []
    for toks in trees:
        edges, tips, nint, stack, i = [], [], 0, [], 0
        expect = True

        def label(i):
            if i < len(toks) and toks[i] not in "(),;" and not toks[i].startswith(":"):
                t = toks[i]
                return (t[1:-1].replace("''", "'") if t.startswith("'") else t), i + 1
            return "", i

        def length(i):
            return i + 1 if i < len(toks) and toks[i].startswith(":") else i

        while True:
            t = toks[i]
            if expect:
                if t == "(":
                    if stack:
                        edges.append((stack[-1], ("n", nint)))
                    stack.append(nint)
                    nint += 1
                    i += 1
                    continue
                lab, i = label(i)
                tips.append(lab)
                if stack:
                    edges.append((stack[-1], ("t", len(tips) - 1)))
                i = length(i)
                expect = False
                if not stack:
                    assert toks[i] == ";"
                    break
            elif t == ",":
                i += 1
                expect = True
            elif t == ")":
                stack.pop()
                _, i = label(i + 1)
                i = length(i)
            elif t == ";":
                if stack:
                    raise ValueError("newick: unbalanced '('")
                break
            else:
                raise ValueError("newick: unexpected " + t)
        n = len(tips)
        if tip_labels is not None:
            pos = {lab: k for k, lab in enumerate(tip_labels)}
            num = [pos[lab] + 1 for lab in tips]
            labels = list(tip_labels)
        else:
            num = list(range(1, n + 1))
            labels = tips
        e = np.array([(n + 1 + a, num[c[1]] if c[0] == "t" else n + 1 + c[1]) for a, c in edges],
                     np.int32).reshape(-1, 2)
        out.append((e, nint, n, labels))
    return out


def tbe_support(ymin, ysize, ntip):
    """Transfer bootstrap expectation (Lemoine et al., Nature 556:452, 2018):
    a reference branch with p tips on its smaller side has, against each
    replicate, transfer index min(ymin, p - 1) (a replicate's tip branch is
    p - 1 away), and TBE = 1 - sum of the indices / ((p - 1) * nrep); NaN when
    p < 2."""
    ym = np.asarray(ymin, dtype=np.int64).reshape(-1, len(ysize))
    out = np.empty(len(ysize))
    for y, sz in enumerate(ysize):
        p = min(int(sz), ntip - int(sz))
        if p < 2 or ym.shape[0] == 0:
            out[y] = np.nan
            continue
        s = sum(min(int(v), p - 1) for v in ym[:, y])
        out[y] = 1.0 - float(s) / (float(p - 1) * float(ym.shape[0]))
    return out
