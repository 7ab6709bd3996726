"""Seeded synthetic inputs for the CasPhy hot path (numpy PCG64, never rand()).

Every generator is deterministic in its seed, so the HIP path, the oracle and
the CPU baseline see byte-identical inputs.  Layouts are the reference's
on-disk ones:

* indel   -- (n, ncol) int32, -1 missing, 0 unmutated, k>0 allele id, alleles
             relabelled per site to 1..k in ascending order of the raw id as
             regularize.mat does (R/CasPhy_functions.R:435-441);
* binary  -- (n, W) uint32 cell-major, site s at bit s&31 of word s>>5
             (src/functions.cpp:54-68);
* DNA     -- (n, nwRaw) uint32, four fasta2bin byte codes per word, base 4w+j in
             byte j (cu/fasta2bin.cu:12-32,189-210);
* bipartition bitsets -- (m, ceil(ntip/32)) uint32, tip t at bit t&31 of word
             t>>5, zero padding (the layout TransferDisStream reads,
             cu/TransferDisStream.cu:141-147).
"""
from __future__ import annotations

import numpy as np

DNA_CODE = {"A": 0x01, "T": 0x02, "C": 0x04, "G": 0x08, "-": 0x10}
_BASES = np.array([0x01, 0x02, 0x04, 0x08], dtype=np.uint8)


def pack_bits(b: np.ndarray) -> np.ndarray:
    """(n, nsite) bool -> (n, ceil(nsite/32)) uint32, LSB first."""
    n, nsite = b.shape
    W = (nsite + 31) // 32
    pad = np.zeros((n, W * 32), dtype=np.uint8)
    pad[:, :nsite] = b
    by = np.packbits(pad, axis=1, bitorder="little")
    return np.ascontiguousarray(by).view("<u4").astype(np.uint32, copy=False).reshape(n, W)


def unpack_bits(p: np.ndarray, nsite: int | None = None) -> np.ndarray:
    p = np.ascontiguousarray(p, dtype="<u4")
    n, W = p.shape
    b = np.unpackbits(p.view(np.uint8).reshape(n, W * 4), axis=1, bitorder="little")
    return b[:, : (W * 32 if nsite is None else nsite)].astype(bool)


def pack_bytes4(codes: np.ndarray) -> np.ndarray:
    """(n, L) uint8 byte codes -> (n, ceil(L/4)) uint32, zero padding bytes."""
    n, L = codes.shape
    nw = (L + 3) // 4
    pad = np.zeros((n, nw * 4), dtype=np.uint8)
    pad[:, :L] = codes
    return np.ascontiguousarray(pad).view("<u4").astype(np.uint32, copy=False).reshape(n, nw)


# --------------------------------------------------------------- trees
def balanced_tree(depth: int):
    """Parent array of a complete binary division tree; node 0 is the root."""
    nn = 2 ** (depth + 1) - 1
    parent = np.full(nn, -1, dtype=np.int64)
    parent[1:] = (np.arange(1, nn) - 1) // 2
    leaves = np.arange(2 ** depth - 1, nn)
    return parent, leaves


def split_tree(n: int, rng: np.random.Generator):
    """Random binary division tree with n leaves grown by splitting a uniformly
    chosen leaf (Yule / coalescent topology).  Returns (parent, leaves, birth)
    where birth[v] is the split time of v's parent lineage under rate-1 Yule."""
    parent = np.full(2 * n - 1, -1, dtype=np.int64)
    birth = np.zeros(2 * n - 1, dtype=np.float64)
    leaves = [0]
    nxt = 1
    t = 0.0
    picks = rng.random(n)
    waits = rng.exponential(1.0, n)
    for k in range(1, n):
        t += waits[k] / k
        i = int(picks[k] * k)
        v = leaves[i]
        a, b = nxt, nxt + 1
        nxt += 2
        parent[a] = v
        parent[b] = v
        birth[a] = birth[b] = t
        leaves[i] = a
        leaves.append(b)
    return parent, np.array(leaves, dtype=np.int64), birth, t


def _levels(parent: np.ndarray):
    depth = np.zeros(parent.size, dtype=np.int64)
    for v in range(1, parent.size):      # parents always precede children
        depth[v] = depth[parent[v]] + 1
    order = np.argsort(depth, kind="stable")
    bounds = np.searchsorted(depth[order], np.arange(depth.max() + 2))
    return [order[bounds[d]:bounds[d + 1]] for d in range(1, depth.max() + 1)]


# --------------------------------------------------------------- configs
def indel_matrix(seed: int = 1, depth: int = 10, nsite: int = 200, p: float = 0.02,
                 nallele: int = 100, zipf_s: float = 1.5, missing: float = 0.1) -> np.ndarray:
    """Indel lineage matrix: balanced tree, irreversible per-generation
    mutation with probability p, Zipf(s) allele ids, then `missing` entries -1."""
    rng = np.random.default_rng(seed)
    parent, leaves = balanced_tree(depth)
    probs = np.arange(1, nallele + 1, dtype=np.float64) ** -zipf_s
    probs /= probs.sum()
    st = np.zeros((parent.size, nsite), dtype=np.int32)
    for lev in _levels(parent):
        s = st[parent[lev]].copy()
        mut = (s == 0) & (rng.random(s.shape) < p)
        s[mut] = rng.choice(nallele, size=int(mut.sum()), p=probs).astype(np.int32) + 1
        st[lev] = s
    mat = st[leaves]
    mat[rng.random(mat.shape) < missing] = -1
    return regularize(mat)


def regularize(mat: np.ndarray) -> np.ndarray:
    """Per-site relabel: -1 -> -1, 0 -> 0, alleles -> rank 1..k (ascending)."""
    out = np.empty_like(mat)
    for s in range(mat.shape[1]):
        col = mat[:, s]
        u = np.unique(np.concatenate([col, [-1, 0]]))
        out[:, s] = np.searchsorted(u, col) - 1
    return out


def indel_weights(mat: np.ndarray) -> np.ndarray:
    """Unit weights, nState x ncol, nState = max alleles per site
    (R/CasPhy_functions.R:583-586)."""
    nstate = max(1, int(mat.max()))
    return np.ones((nstate, mat.shape[1]), dtype=np.float64)


def binary_matrix(seed: int = 2, n: int = 10000, nsite: int = 4096, p: float = 0.005,
                  missing: float = 0.05) -> np.ndarray:
    """Base-editor matrix packed to (n, nsite/32) uint32."""
    rng = np.random.default_rng(seed)
    parent, leaves, _, _ = split_tree(n, rng)
    st = np.zeros((parent.size, nsite), dtype=bool)
    for lev in _levels(parent):
        st[lev] = st[parent[lev]] | (rng.random((lev.size, nsite)) < p)
    b = st[leaves]
    b &= rng.random(b.shape) >= missing
    return pack_bits(b)


def dna_matrix(seed: int = 3, n: int = 100000, ncol: int = 1000, height: float = 0.3,
               gap: float = 0.02, chunk: int = 8192) -> np.ndarray:
    """JC69 alignment on a rate-1 Yule tree scaled to mean root-to-tip
    `height` substitutions/site, `gap` random gaps, as fasta2bin byte codes."""
    rng = np.random.default_rng(seed)
    parent, leaves, birth, _ = split_tree(n, rng)
    # a tip's lineage ends when the process stops: give tips the mean depth
    depth_time = birth.copy()
    tip_end = birth[leaves].max()
    end = np.empty_like(birth)
    end[:] = np.nan
    children_end = np.full(parent.size, np.inf)
    for v in range(1, parent.size):
        pv = parent[v]
        children_end[pv] = min(children_end[pv], birth[v])
    end = np.where(np.isinf(children_end), tip_end, children_end)
    blen = end - depth_time
    blen[0] = 0.0
    scale = height / max(tip_end, 1e-12)
    blen = np.maximum(blen, 0.0) * scale
    pchg = 0.75 * (1.0 - np.exp(-4.0 / 3.0 * blen))
    st = np.empty((parent.size, ncol), dtype=np.uint8)
    st[0] = rng.integers(0, 4, ncol, dtype=np.uint8)
    for lev in _levels(parent):
        s = st[parent[lev]].copy()
        chg = rng.random(s.shape) < pchg[lev, None]
        s[chg] = (s[chg] + rng.integers(1, 4, int(chg.sum()), dtype=np.uint8)) % 4
        st[lev] = s
    codes = _BASES[st[leaves]]
    codes[rng.random(codes.shape) < gap] = 0x10
    return pack_bytes4(codes)


def query_rows(seed: int, n: int, R: int) -> np.ndarray:
    """tryRadius query rows (the reference draws them with time-seeded rand(),
    cu/tryRadius.cu:121,129; here they are an explicit seeded list)."""
    return np.random.default_rng(seed).integers(0, n, R, dtype=np.int64).astype(np.uint32)


# --------------------------------------------------------------- bipartitions
def yule_parent(n: int, rng: np.random.Generator):
    parent, leaves, _, _ = split_tree(n, rng)
    return parent, leaves


def spr(parent: np.ndarray, leaves: np.ndarray, moves: int, rng: np.random.Generator) -> np.ndarray:
    """`moves` random subtree-prune-and-regraft moves on a rooted binary tree
    given as a parent array; node ids and tip ids are kept."""
    par = parent.copy()
    nn = par.size
    kids = [[] for _ in range(nn)]
    for v in range(1, nn):
        kids[par[v]].append(v)
    root = 0
    done = 0
    while done < moves:
        s = int(rng.integers(1, nn))
        p = par[s]
        if p == root:
            continue
        t = int(rng.integers(0, nn))
        # target edge (par[t], t) must lie outside the pruned subtree and off p
        a = t
        bad = False
        while a != -1:
            if a == s:
                bad = True
                break
            a = par[a]
        if bad or t == p or t == root or par[t] == p:
            continue
        g = par[p]
        sib = kids[p][0] if kids[p][1] == s else kids[p][1]
        # remove p: sibling hangs from the grandparent
        kids[g][kids[g].index(p)] = sib
        par[sib] = g
        # reuse p on edge (par[t], t)
        q = par[t]
        kids[q][kids[q].index(t)] = p
        par[p] = q
        kids[p] = [t, s]
        par[t] = p
        done += 1
    return par


def clade_bitsets(parent: np.ndarray, leaves: np.ndarray, ntip: int) -> np.ndarray:
    """Descendant-tip bitsets of every internal non-root node, (ntip-2, W),
    ordered by node id.  Built level by level from the deepest nodes up."""
    nn = parent.size
    W = (ntip + 31) // 32
    kids = np.full((nn, 2), -1, dtype=np.int64)
    order = np.argsort(parent[1:], kind="stable") + 1
    ps = parent[order]
    first = np.ones(order.size, bool)
    first[1:] = ps[1:] != ps[:-1]
    kids[ps[first], 0] = order[first]
    kids[ps[~first], 1] = order[~first]
    bits = np.zeros((nn, W), dtype=np.uint32)
    t = np.arange(ntip)
    bits[leaves, t >> 5] = np.uint32(1) << (t & 31).astype(np.uint32)
    levels = _levels(parent)
    internal_mask = kids[:, 0] >= 0
    for lev in reversed([np.array([0])] + levels):
        iv = lev[internal_mask[lev]]
        if iv.size:
            bits[iv] = bits[kids[iv, 0]] | bits[kids[iv, 1]]
    internal = np.nonzero(internal_mask)[0]
    internal = internal[internal != 0]
    return np.ascontiguousarray(bits[internal])


def ape_edges(parent: np.ndarray, leaves: np.ndarray):
    """R's phylo$edge for a rooted tree given as a parent array (root 0) with
    tip t at node leaves[t]: 1-based, tips 1..ntip, internal nodes numbered
    from ntip + 1 in preorder, edges in cladewise order (depth first, children
    in increasing node id).  Returns (edge (nedge, 2) int32, nnode, ntip)."""
    nn = parent.size
    ntip = leaves.size
    kids = [[] for _ in range(nn)]
    for v in range(1, nn):
        kids[int(parent[v])].append(v)
    label = np.zeros(nn, dtype=np.int64)
    label[leaves] = np.arange(1, ntip + 1)
    nxt = ntip + 1
    label[0] = nxt
    nxt += 1
    edges = []
    stack = [(0, 0)]
    while stack:
        v, k = stack.pop()
        if k < len(kids[v]):
            stack.append((v, k + 1))
            c = kids[v][k]
            if kids[c]:
                label[c] = nxt
                nxt += 1
            edges.append((label[v], label[c]))
            if kids[c]:
                stack.append((c, 0))
    return np.array(edges, dtype=np.int32).reshape(-1, 2), int(nxt - ntip - 1), ntip


def random_tree(rng: np.random.Generator, ntip: int, nint: int):
    """Random rooted tree in ape's edge order: `nint` internal nodes (a random
    recursive tree, so some have a single child), tips hung on random
    internal nodes.  Returns (edge (nedge, 2) int32 1-based, nnode)."""
    nall = nint + ntip
    parent = np.full(nall, -1, np.int64)
    for v in range(1, nint):
        parent[v] = rng.integers(0, v)
    parent[nint:] = rng.integers(0, nint, ntip)
    # every internal node needs a child: childless ones take a tip of a node
    # with at least two children
    kids = np.bincount(parent[1:], minlength=nall)
    for v in range(nint):
        if kids[v] == 0:
            donors = [t for t in range(nint, nall) if kids[parent[t]] >= 2]
            t = donors[int(rng.integers(0, len(donors)))]
            kids[parent[t]] -= 1
            parent[t] = v
            kids[v] += 1
    leaves = rng.permutation(np.arange(nint, nall))
    edge, nnode, nt = ape_edges(parent, leaves)
    assert nt == ntip
    return edge, nnode


def caterpillar(ntip: int):
    """((((1,2),3),4)...): the deepest binary tree."""
    parent = np.full(2 * ntip - 1, -1, np.int64)
    # internal nodes 0..ntip-2 in a chain, node k+1 child of k; tips hang off
    for k in range(1, ntip - 1):
        parent[k] = k - 1
    tips = np.arange(ntip - 1, 2 * ntip - 1)
    parent[tips[:ntip - 1]] = np.arange(ntip - 1)
    parent[tips[-1]] = ntip - 2
    edge, nnode, _ = ape_edges(parent, tips)
    return edge, nnode


_NATIVE = None


def _native():
    """The native tree generator (libcasphy_synth.so, built beside the HIP
    library); raises when it has not been built."""
    global _NATIVE
    if _NATIVE is None:
        import ctypes as C
        import os
        lib = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcasphy_synth.so"))
        lib.casphy_synth_spr_edges.restype = C.c_int64
        lib.casphy_synth_spr_edges.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64, C.c_uint64,
                                               C.c_void_p]
        _NATIVE = lib
    return _NATIVE


def spr_edges_batch(parent: np.ndarray, leaves: np.ndarray, moves: int, nrep: int, seed: int, threads: int = 8):
    """nrep bootstrap trees of a rooted binary tree, each `moves` random SPR
    moves away from it (the moves of spr, drawn by a seeded native generator,
    replicate b with seed (seed, b)), as R edge matrices numbered as
    ape_edges numbers them: [(edge (nn - 1, 2) int32, nnode)]."""
    from concurrent.futures import ThreadPoolExecutor
    lib = _native()
    par = np.ascontiguousarray(parent, dtype=np.int64)
    lv = np.ascontiguousarray(leaves, dtype=np.int64)
    nn = par.size

    def one(b):
        e = np.empty((nn - 1, 2), np.int32)
        nnode = lib.casphy_synth_spr_edges(par.ctypes.data, nn, lv.ctypes.data, lv.size, int(moves),
                                           (int(seed) << 32) ^ int(b), e.ctypes.data)
        if nnode < 0:
            raise ValueError("spr_edges_batch: the tree is not a rooted binary parent array")
        return e, int(nnode)
    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(one, range(nrep)))


# --------------------------------------------------------------- root votes
def vote_inputs(seed: int, ntip: int, nsample: int, ncol: int, mat: np.ndarray | None = None):
    """voteRoot2 inputs: a Yule tree on the first ntip rows of a JC69 alignment
    (fasta2bin byte codes; `mat` or dna_matrix(seed, ntip + nsample, ncol)),
    internal states by a Fitch pass over the byte codes (per position the
    children's common bases, else their union), and the next nsample rows as
    outgroup samples.  Returns (node_state (ntip + nnode, nbin) u32, edge
    (nedge, 2) u32 1-based, mat, idx (nsample,) u32)."""
    if mat is None:
        mat = dna_matrix(seed=seed, n=ntip + nsample, ncol=ncol)
    rng = np.random.default_rng(seed + 1)
    parent, leaves = yule_parent(ntip, rng)
    edge, nnode, ntip = ape_edges(parent, leaves)
    nbin = mat.shape[1]
    st = np.zeros((ntip + nnode, nbin), np.uint32)
    st[:ntip] = mat[:ntip]
    inter = np.zeros((ntip + nnode, nbin * 4), np.uint8)
    union = np.zeros((ntip + nnode, nbin * 4), np.uint8)
    seen = np.zeros(ntip + nnode, bool)
    for a, d in edge[::-1]:  # children before parents
        a, d = int(a) - 1, int(d) - 1
        if d >= ntip:  # internal child complete: its Fitch set
            i, u = inter[d], union[d]
            st[d] = np.where((i & 0x0F) != 0, i, u).view(np.uint32)
        b = st[d].view(np.uint8)
        if seen[a]:
            inter[a] &= b
            union[a] |= b
        else:
            inter[a], union[a], seen[a] = b, b, True
    r = ntip  # the root
    st[r] = np.where((inter[r] & 0x0F) != 0, inter[r], union[r]).view(np.uint32)
    idx = np.arange(ntip, ntip + nsample, dtype=np.uint32)
    return st, edge.astype(np.uint32), mat, idx


def newick(parent: np.ndarray, leaves: np.ndarray, labels=None, rng: np.random.Generator | None = None) -> str:
    """Newick text of a rooted tree given as a parent array (root 0) with tip
    t at node leaves[t] (label labels[t], default "t<t>"), children in
    increasing node id (the order ape_edges lists them); with rng, random
    branch lengths and internal node labels that a reader drops."""
    nn = parent.size
    kids = [[] for _ in range(nn)]
    for v in range(1, nn):
        kids[int(parent[v])].append(v)
    name = {}
    for t, v in enumerate(leaves):
        name[int(v)] = labels[t] if labels is not None else f"t{t}"
    out = []
    stack = [(0, 0)]
    while stack:
        v, k = stack.pop()
        if not kids[v]:
            out.append(name[v])
        elif k == 0:
            out.append("(")
            stack.append((v, 1))
            stack.append((kids[v][0], 0))
            continue
        elif k < len(kids[v]):
            out.append(",")
            stack.append((v, k + 1))
            stack.append((kids[v][k], 0))
            continue
        else:
            out.append(")")
            if rng is not None:
                out.append(f"n{v}")
        if rng is not None and v != 0:
            out.append(f":{rng.random():.5f}")
    return "".join(out) + ";"


# --------------------------------------------------------------- serving mix
def serving_jobs(seed: int, njobs: int, min_cells: int, max_cells: int, nbin_rows: int, nindel_rows: int,
                 tbe_every: int = 4, nrep: int = 3):
    """The serving mix: job sizes log-uniform in [min_cells, max_cells]; every
    tbe_every-th job is a transfer-distance job on a Yule tree of that many
    tips against nrep replicates 5% SPR moves away (as R edge matrices), the
    others alternate binary and indel tree builds over that many rows, drawn
    without replacement and ascending, of resident matrices of nbin_rows /
    nindel_rows rows.  Returns [(kind, data)]: ("bit" | "indel", rows) or
    ("trees", (ref_edge, ref_nnode, rep_edges, rep_nnodes, ntip))."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(njobs):
        n = int(round(np.exp(rng.uniform(np.log(min_cells), np.log(max_cells)))))
        if tbe_every and k % tbe_every == tbe_every - 1:
            parent, leaves = yule_parent(max(n, 4), rng)
            ref_e, ref_n, ntip = ape_edges(parent, leaves)
            reps = spr_edges_batch(parent, leaves, int(np.ceil(0.05 * (ntip - 3))), nrep, seed * 100003 + k)
            jobs.append(("trees", (ref_e, ref_n, [e for e, _ in reps], [m for _, m in reps], ntip)))
        else:
            bit = len(jobs) % 2 == 0
            pool = nbin_rows if bit else nindel_rows
            rows = np.sort(rng.choice(pool, min(n, pool), replace=False)).astype(np.uint32)
            jobs.append(("bit" if bit else "indel", rows))
    return jobs
