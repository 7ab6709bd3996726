"""Host-side mirror of CasPhy's hot-path interface over the HIP library.

Two layers, both backed only by libcasphy_hip.so (no CPU fallback):

1. The Rcpp twins, with the reference's names and argument meaning
   (R/RcppExports.R; src/functions.cpp):
     greedyJoinEdgeBitReverse(mat, weight, verbose, ancestralType)
     greedyJoinEdge2(mat, weight, tipSize, minSize, minRatio, verbose)
     radiusWeightedParallel(i, mat, maxChar, weight, radius)
     radiusWeightedDNAParallel(i, mat, weight, radius)
   Matrices are numpy arrays holding the same bytes as the R matrices: an R
   matrix with cells as columns (features x cells) is a C-order (n, W) array,
   one with cells as rows (cells x words) is a C-order (W, n) array.

2. The executables' computations (cu/*.cu main functions) on in-memory arrays:
     join_edge_bit_reverse, join_edge_indel, radius_neighbor,
     radius_neighbor_dna, try_radius, try_radius_dna, transfer_dis.

Errors raise CasphyError (the status code and message of the C ABI).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import CasphyError, check, load

__all__ = [
    "CasphyError", "score_bit_matrix", "score_indel_matrix",
    "radius_neighbor", "radius_neighbor_dna", "try_radius", "try_radius_dna",
    "try_radii", "try_radii_dna", "join_bit", "join_indel", "cladewise",
    "join_bit_batch", "join_indel_batch", "join_indel_bootstrap", "transfer_dis", "transfer_dis_batch", "join_edge_bit_reverse", "join_edge_indel",
    "greedyJoinEdgeBitReverse", "greedyJoinEdge2", "radiusWeightedParallel",
    "radiusWeightedDNAParallel", "mfma_tuning", "tip_merge", "node_keep", "clade_bitsets", "transfer_trees",
    "vote_root", "find_root", "findRootAVX2Big", "fasta_encode", "fasta2bin", "read_newick", "tbe_support",
    "transfer_support",
]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _take_list(ptr, total) -> np.ndarray:
    L = load()
    try:
        if total == 0:
            return np.zeros(0, dtype=np.uint32)
        buf = np.ctypeslib.as_array(ptr, shape=(total,))
        return buf.copy()
    finally:
        L.casphy_free(C.cast(ptr, C.c_void_p))


# ------------------------------------------------------------------ scores
def score_bit_matrix(mat_cellmajor, weight=None) -> np.ndarray:
    """n x n int32 shared-mutation scores of bit-packed cells (diagonal = self)."""
    m = _u32(mat_cellmajor)
    n, W = m.shape
    w = _f64(weight)
    S = np.empty((n, n), dtype=np.int32)
    check(load().casphy_score_bit_matrix(_p(m), n, W, _p(w), _p(S)))
    return S


def score_indel_matrix(mat_cellmajor, weight=None) -> np.ndarray:
    """n x n int32 indel scores; weight is (nState, ncol) state-major or None."""
    m = _i32(mat_cellmajor)
    n, ncol = m.shape
    w = _f64(weight)
    nstate = 0 if w is None else w.shape[0]
    S = np.empty((n, n), dtype=np.int32)
    check(load().casphy_score_indel_matrix(_p(m), n, ncol, _p(w), nstate, _p(S)))
    return S


# ------------------------------------------------------------------ radius
def radius_neighbor(bmat_wordmajor, radius: int, weight=None):
    """(nNeighbor, idx) of radiusNeighbor: .bmat payload (W, n) in, the
    .nNeighbor counts and the concatenated ascending .idx lists out."""
    m = _u32(bmat_wordmajor)
    W, n = m.shape
    w = _f64(weight)
    counts = np.empty(n, dtype=np.uint32)
    ptr = C.POINTER(C.c_uint32)()
    total = C.c_uint64(0)
    check(load().casphy_radius_bit(_p(m), W, n, _p(w), int(radius), _p(counts), C.byref(ptr), C.byref(total)))
    return counts, _take_list(ptr, total.value)


def radius_neighbor_dna(tbmat, radius: int):
    """(nNeighbor, idx) of radiusNeighborDNA on fasta2bin codes (n, nwRaw)."""
    m = _u32(tbmat)
    n, nw = m.shape
    counts = np.empty(n, dtype=np.uint32)
    ptr = C.POINTER(C.c_uint32)()
    total = C.c_uint64(0)
    check(load().casphy_radius_dna(_p(m), n, nw, int(radius), _p(counts), C.byref(ptr), C.byref(total)))
    return counts, _take_list(ptr, total.value)


def try_radii(bmat_wordmajor, radii, q, weight=None) -> np.ndarray:
    """Counts (len(radii), R) of tryRadius for every radius in one launch."""
    m = _u32(bmat_wordmajor)
    W, n = m.shape
    r = np.ascontiguousarray(radii, dtype=np.int32)
    q = _u32(q)
    w = _f64(weight)
    out = np.zeros((r.size, q.size), dtype=np.uint32)
    check(load().casphy_try_radii_bit(_p(m), W, n, _p(w), _p(r), r.size, _p(q), q.size, _p(out)))
    return out


def try_radius(bmat_wordmajor, radius: int, q, weight=None) -> np.ndarray:
    return try_radii(bmat_wordmajor, [radius], q, weight)[0]


def try_radii_dna(tbmat, radii, q) -> np.ndarray:
    m = _u32(tbmat)
    n, nw = m.shape
    r = np.ascontiguousarray(radii, dtype=np.int32)
    q = _u32(q)
    out = np.zeros((r.size, q.size), dtype=np.uint32)
    check(load().casphy_try_radii_dna(_p(m), n, nw, _p(r), r.size, _p(q), q.size, _p(out)))
    return out


def try_radius_dna(tbmat, radius: int, q) -> np.ndarray:
    return try_radii_dna(tbmat, [radius], q)[0]


# ------------------------------------------------------------------ joins
def join_bit(mat_cellmajor, weight=None, ancestral_type: int = 1):
    """(edge (2n-2, 2) raw join order, node_state (2n-1, W), node_bistate)
    indexed by node id; node_state is seq & ~bistate."""
    m = _u32(mat_cellmajor)
    n, W = m.shape
    w = _f64(weight)
    edge = np.empty((2 * n - 2, 2), dtype=np.uint32)
    ns = np.empty((2 * n - 1, W), dtype=np.uint32)
    nb = np.empty((2 * n - 1, W), dtype=np.uint32)
    check(load().casphy_join_bit(_p(m), n, W, _p(w), int(ancestral_type), _p(edge), _p(ns), _p(nb)))
    return edge, ns, nb


def join_indel(mat_cellmajor, weight=None, tip_size=None, min_size: int = 0, min_ratio: float = 2.0):
    """(edge (2n-2, 2), node_state (2n-1, ncol) int32 by node id, branch (2n-2,))."""
    m = _i32(mat_cellmajor)
    n, ncol = m.shape
    w = _f64(weight)
    nstate = 0 if w is None else w.shape[0]
    ts = None if tip_size is None else _u32(tip_size)
    edge = np.empty((2 * n - 2, 2), dtype=np.uint32)
    ns = np.empty((2 * n - 1, ncol), dtype=np.int32)
    br = np.empty(2 * n - 2, dtype=np.uint32)
    check(load().casphy_join_indel(_p(m), n, ncol, _p(w), nstate, _p(ts), int(min_size), float(min_ratio),
                                   _p(edge), _p(ns), _p(br)))
    return edge, ns, br


def cladewise(edge_raw, n: int):
    """(edge_out (2n-2, 2) relabelled in DFS order, internal_order (n-1,))."""
    e = _u32(edge_raw)
    eo = np.empty((2 * n - 2, 2), dtype=np.uint32)
    io = np.empty(n - 1, dtype=np.uint32)
    check(load().casphy_cladewise(_p(e), n, _p(eo), _p(io)))
    return eo, io


def join_bit_batch(mats, weight=None, ancestral_type: int = 1):
    """Independent binary joins in one launch; returns a list of raw edges."""
    mats = [_u32(m) for m in mats]
    W = mats[0].shape[1]
    sizes = [m.shape[0] for m in mats]
    offsets = np.zeros(len(mats) + 1, dtype=np.uint32)
    offsets[1:] = np.cumsum(sizes)
    allm = np.ascontiguousarray(np.concatenate(mats, axis=0))
    w = _f64(weight)
    edges = np.empty((int(sum(2 * s - 2 for s in sizes)), 2), dtype=np.uint32)
    check(load().casphy_join_bit_batch(_p(allm), _p(offsets), len(mats), W, _p(w), int(ancestral_type),
                                       _p(edges)))
    out, pos = [], 0
    for s in sizes:
        out.append(edges[pos:pos + 2 * s - 2])
        pos += 2 * s - 2
    return out


def join_indel_batch(mats, weight=None, tip_sizes=None, min_size: int = 0, min_ratio: float = 2.0):
    """Independent indel joins (same site count) in one persistent launch."""
    mats = [_i32(m) for m in mats]
    ncol = mats[0].shape[1]
    sizes = [m.shape[0] for m in mats]
    offsets = np.zeros(len(mats) + 1, dtype=np.uint32)
    offsets[1:] = np.cumsum(sizes)
    allm = np.ascontiguousarray(np.concatenate(mats, axis=0))
    w = _f64(weight)
    nstate = 0 if w is None else w.shape[0]
    ts = None if tip_sizes is None else _u32(np.concatenate(tip_sizes))
    edges = np.empty((int(sum(2 * s - 2 for s in sizes)), 2), dtype=np.uint32)
    check(load().casphy_join_indel_batch(_p(allm), _p(offsets), len(mats), ncol, _p(w), nstate, _p(ts),
                                         int(min_size), float(min_ratio), _p(edges)))
    out, pos = [], 0
    for s in sizes:
        out.append(edges[pos:pos + 2 * s - 2])
        pos += 2 * s - 2
    return out


def join_indel_bootstrap(mat_cellmajor, cols):
    """Column-resample replicates of one indel matrix joined in one batched
    launch: replicate r is mat[:, cols[r]] (cols: nrep x ncol indices), built
    on the device; returns one raw edge array per replicate, each equal to
    join_indel(mat[:, cols[r]])[0] with unit weights."""
    m = _i32(mat_cellmajor)
    n, ncol = m.shape
    c = _u32(np.asarray(cols).reshape(-1, ncol))
    nrep = c.shape[0]
    edges = np.empty((nrep * (2 * n - 2), 2), dtype=np.uint32)
    check(load().casphy_join_indel_bootstrap(_p(m), n, ncol, _p(c), nrep, _p(edges)))
    return [edges[r * (2 * n - 2):(r + 1) * (2 * n - 2)] for r in range(nrep)]


def join_edge_bit_reverse(tbmat, weight=None, ancestral_type: int = 1):
    """The joinEdgeBitReverse executable: .tbmat payload (n, W) in,
    (.edge2 payload (2n-2, 2), .nodeState2 payload (n-1, W)) out."""
    m = _u32(tbmat)
    n = m.shape[0]
    edge, ns, _ = join_bit(m, weight, ancestral_type)
    eo, order = cladewise(edge, n)
    return eo, ns[order]


def join_edge_indel(tbmat_shifted, weight, n_state: int, tip_size, min_size: int, min_ratio: float):
    """The joinEdgeIndel executable: .tbmat payload with states +1, the .w
    weights (nState*ncol, state-major), nState, .tipSize, minSize, minRatio ->
    (.edge2 payload, .nodeState2 payload with states +1)."""
    m = np.ascontiguousarray(tbmat_shifted, dtype=np.int64).astype(np.int32) - 1
    ncol = m.shape[1]
    w = None if weight is None else np.asarray(weight, dtype=np.float64).reshape(n_state, ncol)
    edge, ns, _ = join_indel(m, w, tip_size, min_size, min_ratio)
    eo, order = cladewise(edge, m.shape[0])
    return eo, (ns[order] + 1).astype(np.uint32)


# ------------------------------------------------------------------ transfer
def transfer_dis(Y, X, ntip: int):
    """(ymin, ysize, xmin, xsize) of TransferDisStream on bitset rows."""
    Y, X = _u32(Y), _u32(X)
    ny, ncol = Y.shape
    nx = X.shape[0]
    if X.shape[1] != ncol:
        raise CasphyError(-1, "X and Y must have the same number of words per bitset")
    ymin = np.empty(ny, np.uint32)
    ysize = np.empty(ny, np.uint32)
    xmin = np.empty(nx, np.uint32)
    xsize = np.empty(nx, np.uint32)
    check(load().casphy_transfer(_p(Y), ny, _p(X), nx, ncol, int(ntip), _p(ymin), _p(ysize), _p(xmin),
                                 _p(xsize)))
    return ymin, ysize, xmin, xsize


def transfer_dis_batch(Y, Xs, ntip: int, counts=None):
    """Transfer distances of reference bitsets Y against several replicate
    sets in one launch: (ymin (nrep, ny), ysize, [xmin_b], [xsize_b]).
    Xs is a list of (rows_b, ncol) arrays, or, with counts = [rows_b], one
    (sum rows_b, ncol) array holding the replicates back to back (no copy)."""
    Y = _u32(Y)
    ny, ncol = Y.shape
    if counts is None:
        Xs = [_u32(x) for x in Xs]
        rows = [x.shape[0] for x in Xs]
        X = np.ascontiguousarray(np.concatenate(Xs, axis=0)) if Xs else np.zeros((0, ncol), np.uint32)
    else:
        rows = [int(c) for c in counts]
        X = _u32(Xs)
        if X.ndim != 2 or X.shape[0] != sum(rows) or (rows and X.shape[1] != ncol):
            raise ValueError("transfer_dis_batch: Xs must be (sum(counts), ncol)")
    offs = np.zeros(len(rows) + 1, dtype=np.uint32)
    offs[1:] = np.cumsum(rows)
    nrep = len(rows)
    ymin = np.empty((nrep, ny), np.uint32)
    ysize = np.empty(ny, np.uint32)
    xmin = np.empty(max(int(offs[-1]), 1), np.uint32)
    xsize = np.empty(max(int(offs[-1]), 1), np.uint32)
    check(load().casphy_transfer_batch(_p(Y), ny, _p(X), _p(offs), nrep, ncol, int(ntip), _p(ymin), _p(ysize),
                                       _p(xmin), _p(xsize)))
    xm = [xmin[offs[b]:offs[b + 1]] for b in range(nrep)]
    xs = [xsize[offs[b]:offs[b + 1]] for b in range(nrep)]
    return ymin, ysize, xm, xs


def tip_merge(nneighbor, idx, p: float) -> np.ndarray:
    """cu/tipMerge.cu on .nNeighbor / .idx contents: int32 label per tip (the
    .member lines)."""
    nnb, ids = _u32(nneighbor), _u32(idx)
    n = nnb.size
    member = np.empty(max(n, 1), np.int32)
    check(load().casphy_tip_merge(_p(nnb), _p(ids) if ids.size else None, n, float(p), _p(member)))
    return member[:n]


def clade_bitsets(edge, nnode: int, ntip: int):
    """nodeDesBig on R's 1-based edge matrix (nedge, 2) in cladewise order:
    (rows (nnode-1, ceil(ntip/32)) u32 in nodeDesBig's order, node sizes as it
    returns them)."""
    e = np.ascontiguousarray(edge, dtype=np.int32)
    anc, des = np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1])
    nbin = ((ntip - 1) >> 5) + 1
    rows = np.empty((max(nnode - 1, 1), nbin), np.uint32)
    sizes = np.empty(max(nnode - 1, 1), np.int32)
    nr = C.c_uint32(0)
    check(load().casphy_clade_bitsets(_p(anc), _p(des), e.shape[0], int(nnode), int(ntip), _p(rows), C.byref(nr),
                                      _p(sizes)))
    return rows[: nr.value], sizes[: max(nnode - 1, 0)]


def transfer_trees(ref_edge, ref_nnode: int, rep_edges, rep_nnodes, ntip: int):
    """Transfer distances of a reference tree against replicate trees, all as
    R 1-based edge matrices (cladewise): (ymin (nrep, ref_nnode-1), ysize,
    [xmin_b], [xsize_b]) over the trees' clade rows in nodeDesBig's order."""
    ye = np.ascontiguousarray(ref_edge, dtype=np.int32)
    ya, yd = np.ascontiguousarray(ye[:, 0]), np.ascontiguousarray(ye[:, 1])
    reps = [np.ascontiguousarray(e, dtype=np.int32) for e in rep_edges]
    nrep = len(reps)
    eoff = np.zeros(nrep + 1, np.uint32)
    eoff[1:] = np.cumsum([e.shape[0] for e in reps])
    allx = np.concatenate(reps) if nrep else np.zeros((0, 2), np.int32)
    xa, xd = np.ascontiguousarray(allx[:, 0]), np.ascontiguousarray(allx[:, 1])
    xn = np.ascontiguousarray(rep_nnodes, dtype=np.uint32)
    ny = int(ref_nnode) - 1
    rows = [int(n) - 1 for n in rep_nnodes]
    roff = np.zeros(nrep + 1, np.int64)
    roff[1:] = np.cumsum(rows)
    ymin = np.empty((nrep, ny), np.uint32)
    ysize = np.empty(ny, np.uint32)
    xmin = np.empty(max(int(roff[-1]), 1), np.uint32)
    xsize = np.empty(max(int(roff[-1]), 1), np.uint32)
    check(load().casphy_transfer_trees(_p(ya), _p(yd), ye.shape[0], int(ref_nnode), _p(xa) if xa.size else None,
                                       _p(xd) if xd.size else None, _p(eoff), _p(xn), nrep, int(ntip), _p(ymin),
                                       _p(ysize), _p(xmin), _p(xsize)))
    return ymin, ysize, [xmin[roff[b]:roff[b + 1]] for b in range(nrep)], [xsize[roff[b]:roff[b + 1]] for b in range(nrep)]


def node_keep(edge, node_state, outmat, cutoff: int, indel: bool = False) -> np.ndarray:
    """nodeKeepBit / nodeKeepIndel on .edge2 / .nodeState2 / .tboutmat
    payloads: edge (nedge, 2) 0-based, node_state (nnode, width), outmat
    (nout, width); the kept internal node ids in edge order (.retainNode)."""
    e, st, om = _u32(edge), _u32(node_state), _u32(outmat)
    nnode, width = st.shape
    nout = om.shape[0] if om.ndim == 2 else 0
    if nout and om.shape[1] != width:
        raise CasphyError(-1, "outmat rows must have the node rows' width")
    kept = np.empty(max(nnode, 1), np.uint32)
    nk = C.c_uint32(0)
    fn = load().casphy_node_keep_indel if indel else load().casphy_node_keep_bit
    check(fn(_p(e) if e.size else None, e.shape[0] if e.ndim == 2 else 0, _p(st) if st.size else None, nnode, width,
             _p(om) if om.size else None, nout, int(cutoff), _p(kept), C.byref(nk)))
    return kept[:nk.value]


# ------------------------------------------------------------------ trees from Newick
def read_newick(text, tip_labels=None):
    """Every tree of a Newick text (str or bytes) as ape's read.tree numbers
    it: a list of (edge (nedge, 2) int32 1-based cladewise, nnode, ntip, tip
    labels).  With tip_labels, tips are numbered by the position of their
    label there (so that trees share one numbering)."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    order = None
    if tip_labels is not None:
        order = b"".join((t.encode() if isinstance(t, str) else bytes(t)) + b"\0" for t in tip_labels)
    buf = C.create_string_buffer(data, len(data) + 1)
    base = C.cast(buf, C.c_void_p).value
    obuf = C.create_string_buffer(order, len(order)) if order else None
    L = load()
    trees, pos = [], 0
    while pos < len(data):
        consumed, ntip, nnode, nedge, nlab = C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        ep, lp = C.c_void_p(), C.c_void_p()
        check(L.casphy_newick_read(base + pos, len(data) - pos, C.cast(obuf, C.c_void_p) if obuf else None,
                                   len(order) if order else 0, C.byref(consumed), C.byref(ntip), C.byref(nnode),
                                   C.byref(ep), C.byref(nedge), C.byref(lp), C.byref(nlab)))
        pos += consumed.value
        if ntip.value == 0:
            break
        try:
            ne = nedge.value
            cols = np.ctypeslib.as_array(C.cast(ep, C.POINTER(C.c_int32)), (2 * ne,)).copy() if ne else \
                np.zeros(0, np.int32)
            labels = C.string_at(lp.value, nlab.value).split(b"\0")[:-1]
        finally:
            L.casphy_free(ep)
            L.casphy_free(lp)
        edge = np.ascontiguousarray(cols.reshape(2, ne).T)
        trees.append((edge, nnode.value, ntip.value, [x.decode("utf-8", "surrogateescape") for x in labels]))
    return trees


def tbe_support(ymin, ysize, ntip: int) -> np.ndarray:
    """Transfer bootstrap expectation of the reference clades from the
    transfer minima (nrep, ny) and clade sizes (ny,) of transfer_trees."""
    ym = _u32(ymin)
    ys = _u32(ysize)
    nrep, ny = (ym.shape if ym.ndim == 2 else (0, ys.size))
    out = np.empty(max(ny, 1), np.float64)
    check(load().casphy_tbe_support(_p(ym) if ym.size else None, nrep, ny, _p(ys) if ys.size else None, int(ntip),
                                    _p(out)))
    return out[:ny]


def transfer_support(ref_tree, rep_trees):
    """TBE of the reference tree's clades (nodeDesBig's row order) over the
    replicate trees, all as read_newick returns them with one tip numbering:
    (support, clade sizes, transfer minima (nrep, ny))."""
    edge, nnode, ntip, _ = ref_tree
    ymin, ysize, _, _ = transfer_trees(edge, nnode, [t[0] for t in rep_trees], [t[1] for t in rep_trees], ntip)
    return tbe_support(ymin, ysize, ntip), ysize, ymin


# ------------------------------------------------------------------ DNA format edge
def _root_vote(fn, node_state, edge, mat, idx, dtype):
    st = _u32(node_state)
    if st.ndim != 2:
        raise CasphyError(-1, "node_state must be (nnode, nbin)")
    nnode, nbin = st.shape
    e = _u32(edge).reshape(-1, 2) if np.size(edge) else np.zeros((0, 2), np.uint32)
    m = _u32(mat).reshape(-1, nbin) if nbin else np.zeros((0, 0), np.uint32)
    nmat = m.shape[0] if nbin else int(np.asarray(mat).shape[0])
    ix = _u32(idx).reshape(-1)
    ns = ix.size
    root, nsh = np.empty(max(ns, 1), dtype), np.empty(max(ns, 1), dtype)
    check(fn(_p(st) if st.size else None, nnode, nbin, _p(e) if e.size else None, e.shape[0],
             _p(m) if m.size else None, nmat, _p(ix) if ns else None, ns, _p(root), _p(nsh)))
    return root[:ns], nsh[:ns]


def vote_root(node_state, edge, mat, idx):
    """voteRoot2 (cu/voteRoot2.cu): node_state (nnode, nbin) u32 byte codes,
    edge (nedge, 2) 1-based (parent, child), samples mat[idx]; per sample the
    0-based node of the most shared positions and that count (.rootVote)."""
    return _root_vote(load().casphy_vote_root, node_state, edge, mat, idx, np.uint32)


def find_root(node_state, edge, mat, idx):
    """findRoot / findRootAVX2Big (src/functions.cpp:4170,4317) on the same
    inputs: per sample the 1-based node (first edge child in edge order among
    the largest counts) and the count, int32."""
    return _root_vote(load().casphy_find_root, node_state, edge, mat, idx, np.int32)


def findRootAVX2Big(edge, nodeState, mat, outIdx):
    """R-shaped twin: edge (nedge, 2) 1-based, nodeState (nPositions, nNodes)
    and mat (nPositions, nmat) column-major as R holds them, outIdx 0-based;
    returns the (nSamples, 2) int32 matrix [rootFound, nShared]."""
    st = np.ascontiguousarray((np.asarray(nodeState).astype(np.int64) & 0xFFFFFFFF).T).astype(np.uint32)
    m = np.ascontiguousarray((np.asarray(mat).astype(np.int64) & 0xFFFFFFFF).T).astype(np.uint32)
    root, nsh = find_root(st, np.asarray(edge, dtype=np.int64).astype(np.uint32), m, np.asarray(outIdx))
    return np.stack([root, nsh], axis=1)


def fasta_encode(seqs):
    """fasta2bin's encoding of equal-length sequences, an (nseq, len) uint8
    array or a list of bytes: (is_variable bool (len,), valid bases (nseq,)
    u32, codes (nseq, width) u32 -- the .bin payload)."""
    if not isinstance(seqs, np.ndarray):
        seqs = [bytes(x) for x in seqs]
        if any(len(x) != len(seqs[0]) for x in seqs):
            raise CasphyError(-1, "All sequences must have the same length for variant filtering")
        seqs = np.frombuffer(b"".join(seqs), np.uint8).reshape(len(seqs), -1) if seqs else np.zeros((0, 0), np.uint8)
    a = np.ascontiguousarray(seqs, dtype=np.uint8)
    nseq, length = a.shape
    isv = np.empty(max(length, 1), np.uint8)
    valid = np.empty(max(nseq, 1), np.uint32)
    out = np.empty(max(nseq * ((length + 3) // 4), 1), np.uint32)
    width = C.c_uint64(0)
    check(load().casphy_fasta_encode(_p(a) if a.size else None, nseq, length, _p(isv), _p(valid), _p(out),
                                     C.byref(width)))
    w = width.value
    return isv[:length].astype(bool), valid[:nseq], out[: nseq * w].reshape(nseq, w)


def fasta2bin(fasta_path: str, prefix: str) -> None:
    """The fasta2bin executable (cu/fasta2bin.cu:256-363) as a call: runs
    casphy_amd/bin/fasta2bin, which writes prefix.bin, prefix_names.txt and
    prefix_positions.txt."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin", "fasta2bin")
    r = subprocess.run([exe, fasta_path, prefix], capture_output=True, text=True)
    if r.returncode != 0:
        raise CasphyError(r.returncode, (r.stderr or r.stdout).strip())


# ------------------------------------------------------------------ Rcpp twins
def greedyJoinEdgeBitReverse(mat, weight, verbose: bool = False, ancestralType: int = 0):
    """src/functions.cpp:3316-3461.  mat: (n, W) bytes of the W x n R matrix.
    Returns dict(edge (2n-2, 2), nodeState (2n-1, W), nodeBistate (2n-1, W))
    with the columns of the R result as rows: entry c < 2n-2 is the child of
    edge c, the last entry is the root."""
    m = _u32(mat)
    n, W = m.shape
    w = None if weight is None else _f64(weight)
    edge, ns, nb = join_bit(m, w, ancestralType)
    order = np.concatenate([edge[:, 1], [2 * n - 2]]).astype(np.int64)
    return {"edge": edge.astype(np.int32), "nodeState": ns[order].astype(np.int32),
            "nodeBistate": nb[order].astype(np.int32)}


def greedyJoinEdge2(mat, weight, tipSize, minSize: int = 0, minRatio: float = 2.0, verbose: bool = False):
    """src/functions.cpp:1880-2017.  mat: (n, ncol) bytes of the ncol x n R
    matrix; weight: (nState, ncol) with weight[a-1, s].  Returns the
    (2n-1, 3+ncol) bytes of the (3+ncol) x (2n-1) R result: per column
    parent, child, branch count, child state; the last column holds the root
    state with zeros in rows 0..2."""
    m = _i32(mat)
    n, ncol = m.shape
    edge, ns, br = join_indel(m, weight, tipSize, minSize, minRatio)
    out = np.zeros((2 * n - 1, 3 + ncol), dtype=np.int32)
    out[:-1, 0:2] = edge
    out[:-1, 2] = br
    out[:-1, 3:] = ns[edge[:, 1].astype(np.int64)]
    out[-1, 3:] = ns[2 * n - 2]
    return out


def _bitmask_result(row_list, n):
    nbin = (n + 31) // 32
    out = np.zeros(nbin + 1, dtype=np.uint32)
    out[0] = row_list.size
    if row_list.size:
        np.bitwise_or.at(out, 1 + (row_list >> 5), (np.uint32(1) << (row_list & 31).astype(np.uint32)))
    return out.view(np.int32)


# R calls the per-row functions once per query row (R/CasPhy_functions.R:
# 160-208, 1730-1780); the lists of every row come from one device pass, kept
# for the next call with the same matrix, weights and radius (one entry).
_LISTS = {}


def _row_lists(kind: str, m: np.ndarray, weight, radius: int):
    import hashlib
    wkey = None if weight is None else hashlib.sha1(np.ascontiguousarray(weight, dtype=np.float64).data).hexdigest()
    key = (kind, m.shape, radius, hashlib.sha1(m.data).hexdigest(), wkey)
    hit = _LISTS.get(key)
    if hit is None:
        counts, idx = radius_neighbor(m, radius, weight) if kind == "bit" else radius_neighbor_dna(m, radius)
        off = np.concatenate([[0], np.cumsum(counts, dtype=np.uint64)])
        _LISTS.clear()
        hit = _LISTS[key] = (off, idx)
    return hit


def radiusWeightedParallel(i: int, mat, maxChar, weight, radius: float):
    """src/functions.cpp:1825-1857.  mat: (W, n) bytes of the n x W R matrix;
    returns IntegerVector(nbin+1) = [count, neighbour bitmask words...].  The
    first call for a matrix computes every row's list; later calls for its
    other rows slice them."""
    m = _u32(mat)
    W, n = m.shape
    r = float(radius)
    if r != int(r):
        raise CasphyError(-1, "radius must be an integer for the exact path")
    off, idx = _row_lists("bit", m, weight, int(r))
    return _bitmask_result(idx[int(off[i]):int(off[i + 1])], n)


def radiusWeightedDNAParallel(i: int, mat, weight, radius: float):
    """src/functions.cpp:4021-4036 with unit weights.  mat: (n, nwRaw); lists
    computed once per matrix and radius, as radiusWeightedParallel."""
    m = _u32(mat)
    n = m.shape[0]
    if weight is not None and not np.all(np.asarray(weight) == 1):
        raise CasphyError(-4, "DNA weights other than 1 are outside the exact path")
    r = float(radius)
    off, idx = _row_lists("dna", m, None, int(np.ceil(r)))
    return _bitmask_result(idx[int(off[i]):int(off[i + 1])], n)


# ------------------------------------------------------------------ resident
def mfma_tuning() -> list:
    """Choices of the tile autotuners so far: one dict per shape class with
    the chosen variant and the time in ms of every candidate -- the
    matrix-core tiles (epilogue, triangle, log2 row counts, sites), then the
    VALU tiles ("form": "valu"; metric, epilogue, triangle, log2 row counts,
    words, weight planes; variants kKCxCB: KC words staged per step, CB
    columns per thread).  CASPHY_MFMA_VARIANT / CASPHY_VALU_VARIANT force a
    variant and CASPHY_AUTOTUNE=0 keeps the defaults."""
    import json
    lib = load()
    n = C.c_size_t(0)
    check(lib.casphy_mfma_tuning(None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(lib.casphy_mfma_tuning(buf, n.value + 1, C.byref(n)))
    return json.loads(buf.value.decode())


class Cells:
    """A packed cell matrix resident in HBM (casphy_cells_* of the C ABI).

    kind "bit": `mat` is the .bmat payload (W, n) (wordmajor=True) or a
    cell-major (n, W) array; kind "dna": fasta2bin codes (n, nwRaw)."""

    def __init__(self, mat, kind: str = "bit", wordmajor: bool = True, weight=None):
        L = load()
        m = _u32(mat)
        if kind == "bit":
            W, n = (m.shape if wordmajor else m.shape[::-1])
            k = 0
        else:
            n, W = m.shape
            k = 1
            wordmajor = False
        self.n = n
        w = _f64(weight)
        h = C.c_void_p()
        check(L.casphy_cells_create(k, _p(m), n, W, int(wordmajor), _p(w), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().casphy_cells_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def radius(self, radius: int) -> int:
        total = C.c_uint64(0)
        check(load().casphy_cells_radius(self._h, int(radius), C.byref(total)))
        self.total = total.value
        return self.total

    def fetch(self):
        counts = np.empty(self.n, np.uint32)
        idx = np.empty(max(self.total, 1), np.uint32)
        check(load().casphy_cells_fetch(self._h, _p(counts), _p(idx)))
        return counts, idx[: self.total]

    def tip_merge(self, p: float) -> np.ndarray:
        """tipMerge labels over the lists of the last radius / autosplit call,
        which stay on the device."""
        member = np.empty(max(self.n, 1), np.int32)
        check(load().casphy_cells_tip_merge(self._h, float(p), _p(member)))
        return member[: self.n]

    def try_radii(self, radii, q) -> np.ndarray:
        r = np.ascontiguousarray(radii, dtype=np.int32)
        q = _u32(q)
        out = np.zeros((r.size, q.size), np.uint32)
        check(load().casphy_cells_try_radii(self._h, _p(r), r.size, _p(q), q.size, _p(out)))
        return out

    def autosplit(self, radii, q, bound: float, counts: bool = False):
        """One autoSplit round on the device: returns (radius, total list
        length) -- radius -1 when no radius qualifies -- plus the tryRadius
        counts (nradii, R) when counts=True.  The lists stay resident (fetch)."""
        r = np.ascontiguousarray(radii, dtype=np.int32)
        q = _u32(q)
        out = np.zeros((r.size, q.size), np.uint32) if counts else None
        rad, total = C.c_int32(0), C.c_uint64(0)
        check(load().casphy_cells_autosplit(self._h, _p(r), r.size, _p(q), q.size, float(bound), C.byref(rad),
                                            C.byref(total), _p(out)))
        self.total = total.value
        return (rad.value, self.total, out) if counts else (rad.value, self.total)

    def split_groups(self, groups, radii, queries, nneighbor_bound: float = 9.5):
        """autoSplit's radius search for many groups of these cells at once
        (casphy_cells_split_groups): groups = list of row-id arrays, queries
        = list of group-local query-id arrays (one per group).  Returns
        (radius per group, -1 where none passes; list length per group); the
        lists stay on the device (group_fetch, group_tip_merge)."""
        G = len(groups)
        goff = np.zeros(G + 1, np.uint32)
        goff[1:] = np.cumsum([len(g) for g in groups])
        qoff = np.zeros(G + 1, np.uint32)
        qoff[1:] = np.cumsum([len(x) for x in queries])
        rows = _u32(np.concatenate([np.asarray(g, np.uint32) for g in groups]))
        q = _u32(np.concatenate([np.asarray(x, np.uint32) for x in queries]))
        r = np.ascontiguousarray(radii, dtype=np.int32)
        rad = np.zeros(G, np.int32)
        nidx = np.zeros(G, np.uint64)
        check(load().casphy_cells_split_groups(self._h, _p(goff), _p(rows), G, _p(r), r.size, _p(qoff), _p(q),
                                               float(nneighbor_bound), _p(rad), _p(nidx)))
        self._group_n = [len(g) for g in groups]
        self._group_total = [int(x) for x in nidx]
        return rad, nidx

    def _group_known(self, g: int):
        if g >= len(getattr(self, "_group_n", [])):  # the library reports it
            d = np.zeros(1, np.uint32)
            check(load().casphy_cells_group_fetch(self._h, int(g), _p(d), _p(d)))

    def group_fetch(self, g: int):
        self._group_known(g)
        n, t = self._group_n[g], self._group_total[g]
        counts = np.empty(n, np.uint32)
        idx = np.empty(max(t, 1), np.uint32)
        check(load().casphy_cells_group_fetch(self._h, int(g), _p(counts), _p(idx)))
        return counts, idx[:t]

    def group_tip_merge(self, g: int, p: float) -> np.ndarray:
        self._group_known(g)
        member = np.empty(max(self._group_n[g], 1), np.int32)
        check(load().casphy_cells_group_tip_merge(self._h, int(g), float(p), _p(member)))
        return member[: self._group_n[g]]

    def last_ms(self):
        a, b = C.c_float(0), C.c_float(0)
        check(load().casphy_cells_last_ms(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value


__all__.append("Cells")


# ------------------------------------------------------------------ serving
class Server:
    """Request loop over one GPU (casphy_server_* of the C ABI): admission
    under an HBM budget without head-of-line blocking, `workers` concurrent
    streams, small joins batched into one persistent launch, packed matrices
    resident across jobs.  submit_* return a job id; wait(job) returns the
    job's output (raw edges for joins, (ymin, ysize, xmin, xsize) for
    transfers) and its latency in ms."""

    def __init__(self, hbm_budget_bytes: int = 64 << 30, workers: int = 4, device: int = 0):
        h = C.c_void_p()
        check(load().casphy_server_create(device, int(hbm_budget_bytes), int(workers), C.byref(h)))
        self._h = h
        self._shape = {}

    def close(self):
        if getattr(self, "_h", None):
            load().casphy_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_join(self, mat, kind: str = "bit", ancestral_type: int = 1) -> int:
        m = _u32(mat) if kind == "bit" else _i32(mat)
        n, width = m.shape
        job = C.c_uint64(0)
        check(load().casphy_server_submit_join(self._h, 0 if kind == "bit" else 2, _p(m), n, width,
                                               int(ancestral_type), C.byref(job)))
        self._shape[job.value] = ("join", n)
        return job.value

    def submit_transfer(self, Y, X, ntip: int) -> int:
        Y, X = _u32(Y), _u32(X)
        job = C.c_uint64(0)
        check(load().casphy_server_submit_transfer(self._h, _p(Y), Y.shape[0], _p(X), X.shape[0], Y.shape[1],
                                                   int(ntip), C.byref(job)))
        self._shape[job.value] = ("transfer", Y.shape[0], X.shape[0])
        return job.value

    def put_matrix(self, mat, kind: str = "bit") -> int:
        """Makes a packed matrix resident on the server's device: n x W u32
        bit rows (kind "bit") or n x ncol int32 states ("indel")."""
        m = _u32(mat) if kind == "bit" else _i32(mat)
        mid = C.c_uint64(0)
        check(load().casphy_server_put_matrix(self._h, 0 if kind == "bit" else 2, _p(m), m.shape[0], m.shape[1],
                                              C.byref(mid)))
        return mid.value

    def drop_matrix(self, mid: int):
        check(load().casphy_server_drop_matrix(self._h, int(mid)))

    def submit_join_rows(self, mid: int, rows, ancestral_type: int = 1) -> int:
        r = _u32(rows)
        job = C.c_uint64(0)
        check(load().casphy_server_submit_join_rows(self._h, int(mid), _p(r), r.size, int(ancestral_type),
                                                    C.byref(job)))
        self._shape[job.value] = ("join", r.size)
        return job.value

    def submit_transfer_trees(self, ref_edge, ref_nnode: int, rep_edges, rep_nnodes, ntip: int) -> int:
        """Transfer distances from trees (R 1-based edge matrices, as
        transfer_trees); wait() returns (ymin (nrep, ny), ysize, [xmin_b], [xsize_b])."""
        ye = np.ascontiguousarray(ref_edge, dtype=np.int32)
        ya, yd = np.ascontiguousarray(ye[:, 0]), np.ascontiguousarray(ye[:, 1])
        reps = [np.ascontiguousarray(e, dtype=np.int32) for e in rep_edges]
        eoff = np.zeros(len(reps) + 1, np.uint32)
        eoff[1:] = np.cumsum([e.shape[0] for e in reps])
        allx = np.concatenate(reps) if reps else np.zeros((0, 2), np.int32)
        xa, xd = np.ascontiguousarray(allx[:, 0]), np.ascontiguousarray(allx[:, 1])
        xn = np.ascontiguousarray(rep_nnodes, dtype=np.uint32)
        job = C.c_uint64(0)
        check(load().casphy_server_submit_transfer_trees(
            self._h, _p(ya), _p(yd), ye.shape[0], int(ref_nnode), _p(xa) if xa.size else None,
            _p(xd) if xd.size else None, _p(eoff), _p(xn), len(reps), int(ntip), C.byref(job)))
        self._shape[job.value] = ("trees", int(ref_nnode) - 1, [int(n) - 1 for n in rep_nnodes])
        return job.value

    def wait(self, job: int):
        kind = self._shape.pop(job)
        if kind[0] == "join":
            size = (2 * kind[1] - 2) * 2
        elif kind[0] == "trees":
            ny, rows = kind[1], kind[2]
            size = len(rows) * ny + ny + 2 * sum(rows)
        else:
            size = 2 * kind[1] + 2 * kind[2]
        out = np.empty(max(size, 1), np.uint32)
        lat = C.c_double(0)
        check(load().casphy_server_wait(self._h, int(job), _p(out), size, C.byref(lat)))
        if kind[0] == "join":
            return out[:size].reshape(-1, 2), lat.value
        if kind[0] == "trees":
            nrep, tot = len(rows), sum(rows)
            ymin = out[:nrep * ny].reshape(nrep, ny)
            ysize = out[nrep * ny:nrep * ny + ny]
            xs = out[nrep * ny + ny:]
            off = np.concatenate([[0], np.cumsum(rows)]).astype(np.int64)
            xmin = [xs[off[b]:off[b + 1]] for b in range(nrep)]
            xsize = [xs[tot + off[b]:tot + off[b + 1]] for b in range(nrep)]
            return (ymin, ysize, xmin, xsize), lat.value
        ny, nx = kind[1], kind[2]
        return (out[:ny], out[ny:2 * ny], out[2 * ny:2 * ny + nx], out[2 * ny + nx:2 * ny + 2 * nx]), lat.value

    def stats(self) -> dict:
        c, l = C.c_uint64(0), C.c_uint64(0)
        p50, p99, jps = C.c_double(0), C.c_double(0), C.c_double(0)
        check(load().casphy_server_stats(self._h, C.byref(c), C.byref(p50), C.byref(p99), C.byref(jps), C.byref(l)))
        return {"completed": c.value, "p50_ms": p50.value, "p99_ms": p99.value, "jobs_per_s": jps.value,
                "launches": l.value}


__all__.append("Server")
