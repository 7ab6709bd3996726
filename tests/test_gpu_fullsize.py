"""Full-size workloads on the GPU, checked through size-independent properties
(the oracle would take minutes here): symmetry and order of radius lists,
self membership, monotonicity in r, agreement between independent kernels
(matrix-core vs VALU tiles, tryRadius vs list lengths, batched vs plain joins),
tree invariants, and sampled rows against the oracle."""
import numpy as np
import pytest

from casphy_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg2():
    return synth.binary_matrix(seed=2, n=10000, nsite=4096, p=0.005, missing=0.05)


def csr(counts, idx):
    off = np.zeros(counts.size + 1, np.int64)
    off[1:] = np.cumsum(counts, dtype=np.int64)
    return off


def test_config2_radius_properties(hip, orc, cfg2, monkeypatch):
    bmat = np.ascontiguousarray(cfg2.T)
    n = cfg2.shape[0]
    pop = synth.unpack_bits(cfg2).sum(1)
    cells = hip.Cells(bmat, "bit", wordmajor=True)
    prev = None
    for r in (4, 8, 16):
        total = cells.radius(r)
        counts, idx = cells.fetch()
        assert idx.size == total == counts.sum()
        off = csr(counts, idx)
        rows = np.repeat(np.arange(n, dtype=np.int64), counts.astype(np.int64))
        # ascending within each list
        d = np.diff(idx.astype(np.int64))
        inner = np.ones(idx.size, bool)
        inner[off[1:-1][off[1:-1] < idx.size]] = False
        assert np.all(d[inner[1:]] > 0)
        # symmetric: (i, j) listed iff (j, i) listed
        a = rows * n + idx
        b = idx.astype(np.int64) * n + rows
        assert np.array_equal(np.sort(a), np.sort(b))
        # self iff popcount >= r
        selfhit = np.zeros(n, bool)
        selfhit[rows[rows == idx]] = True
        assert np.array_equal(selfhit, pop >= r)
        if prev is not None:
            assert np.all(counts <= prev)
        prev = counts
        # sampled rows against the oracle
        for i0 in (0, 4321, n - 8):
            c2, i2 = orc.radius_bit(cfg2, r, None, (i0, i0 + 8))
            assert np.array_equal(counts[i0:i0 + 8], c2)
            assert np.array_equal(idx[off[i0]:off[i0 + 8]], i2)
        # the VALU tiles and every matrix-core variant give the same lists
        monkeypatch.setenv("CASPHY_MFMA", "0")
        c3, i3 = hip.radius_neighbor(bmat, r)
        monkeypatch.setenv("CASPHY_MFMA", "1")
        assert np.array_equal(c3, counts) and np.array_equal(i3, idx)
        for variant in ("64x2", "128x1", "fp4", "fp4b", "fp4d", "fp4w", "fp4x", "fp4y", "wide", "wide8", "reg"):
            monkeypatch.setenv("CASPHY_MFMA_VARIANT", variant)
            c3, i3 = hip.radius_neighbor(bmat, r)
            assert np.array_equal(c3, counts) and np.array_equal(i3, idx), variant
        monkeypatch.delenv("CASPHY_MFMA_VARIANT")
    q = synth.query_rows(2, n, 1024)
    tr = cells.try_radii(np.arange(1, 17), q)
    total = cells.radius(16)
    counts16, _ = cells.fetch()
    assert np.array_equal(tr[15], counts16[q])
    assert np.all(np.diff(tr.astype(np.int64), axis=0) <= 0)


def test_config2_join_invariants(hip, cfg2):
    m = cfg2[:3000]
    n = m.shape[0]
    e, ns, nb = hip.join_bit(m)
    assert sorted(e[:, 1].tolist()) == list(range(2 * n - 2))
    assert np.all(e[0::2, 0] == np.arange(n, 2 * n - 1))
    assert np.all(e[0::2, 1] < e[1::2, 1])
    assert np.array_equal(ns[:n], m)
    # every node state is a subset of the union of its children's seq under type 1
    # and the batched persistent kernel reproduces the same tree
    eb = hip.join_bit_batch([m])[0]
    assert np.array_equal(eb, e)


def test_config1_bootstrap_batch(hip, orc):
    base = synth.indel_matrix(seed=1)
    rng = np.random.default_rng(9)
    reps = [base[:, rng.integers(0, 200, 200)] for _ in range(32)]
    batch = hip.join_indel_batch(reps)
    for k in (0, 17, 31):
        w = np.ones((max(1, reps[k].max()), 200), np.float32)
        e2, _, _ = orc.join_indel(reps[k], w, lazy=True)
        assert np.array_equal(batch[k], e2)
    for k in (1, 2):
        e1, _, _ = hip.join_indel(reps[k])
        assert np.array_equal(batch[k], e1)


def test_config3_dna_sampled(hip, orc):
    mat = synth.dna_matrix(seed=3, n=20000, ncol=1000)
    cells = hip.Cells(mat, "dna")
    total = cells.radius(850)
    counts, idx = cells.fetch()
    off = csr(counts, idx)
    for i0 in (0, 12345, 19990):
        c2, i2 = orc.radius_dna(mat, 850, (i0, i0 + 10))
        assert np.array_equal(counts[i0:i0 + 10], c2)
        assert np.array_equal(idx[off[i0]:off[i0 + 10]], i2)


def test_tbe_large_against_brute_rows(hip, orc):
    rng = np.random.default_rng(4)
    ntip = 8192
    parent, leaves = synth.yule_parent(ntip, rng)
    ref = synth.clade_bitsets(parent, leaves, ntip)
    boot = synth.clade_bitsets(synth.spr(parent, leaves, 400, rng), leaves, ntip)
    ymin, ysize, xmin, xsize = hip.transfer_dis(ref, boot, ntip)
    sel = rng.integers(0, ref.shape[0], 12)
    y2, ys2, _, _ = orc.transfer(ref[sel], boot, ntip)
    assert np.array_equal(ymin[sel], y2) and np.array_equal(ysize[sel], ys2)
    sel = rng.integers(0, boot.shape[0], 12)
    x2, xs2, _, _ = orc.transfer(boot[sel], ref, ntip)
    assert np.array_equal(xmin[sel], x2) and np.array_equal(xsize[sel], xs2)
    same = hip.transfer_dis(ref, ref, ntip)
    assert np.all(same[0] == 0) and np.all(same[2] == 0)


@pytest.mark.timeout(900)
def test_config3_dna_100k_lists_past_2_31(hip, orc, monkeypatch):
    """The DNA configuration at full size (100,000 sequences x 1,000 columns):
    the autoSplit search on the device picks the host rule's radius, and its
    lists hold more than 2^31 ids (u64 offsets, staged read-back).  Checked:
    counts sum to the list length, every list ascending with its own row in
    it, symmetry and tryRadius counts on sampled rows, sampled rows (first,
    middle and last, whose ids sit past 2^31) against the oracle, and the
    VALU 2-bit form giving the same counts and id checksum as the one-hot
    matrix-core form."""
    monkeypatch.delenv("CASPHY_DNA_FORM", raising=False)
    n = 100000
    dna = synth.dna_matrix(seed=3, n=n, ncol=1000)
    cells = hip.Cells(dna, "dna")
    q = synth.query_rows(3, n, 1024)
    bound = 9.5 - np.log10(n)
    r, total = -1, 0
    for start in range(100, 1000, 160):
        radii = np.arange(start, start + 160, 10)
        cnt = cells.try_radii(radii, q)
        passing = [int(radii[k]) for k in range(len(radii)) if np.log10(max(cnt[k].mean(), 1e-30)) <= bound]
        r_dev, t_dev = cells.autosplit(radii, q, bound)
        if passing:
            r = passing[0]
            assert r_dev == r
            total = t_dev
            break
        assert r_dev == -1 and t_dev == 0
    assert r > 0 and total > 2 ** 31
    counts, idx = cells.fetch()
    assert idx.size == total == int(counts.sum(dtype=np.uint64))
    off = csr(counts, idx)
    assert off[-1] == total
    # ascending within every list, own row present (>= 640 non-gap positions per row)
    step = 10000
    for i0 in range(0, n, step):
        i1 = min(n, i0 + step)
        seg = idx[off[i0]:off[i1]]
        up = seg[1:] > seg[:-1]
        starts = (off[i0 + 1:i1] - off[i0] - 1).astype(np.int64)  # last id of a row, then the next row
        up[starts[(starts >= 0) & (starts < up.size)]] = True
        assert up.all()
        rows = np.repeat(np.arange(i0, i1, dtype=np.uint32), counts[i0:i1].astype(np.int64))
        assert np.count_nonzero(seg == rows) == i1 - i0
        del seg, up, rows
    rng = np.random.default_rng(33)
    sample = np.concatenate([[0, n // 2, n - 1], rng.integers(0, n, 5)])
    for i in sample:
        lst = idx[off[i]:off[i + 1]]
        # symmetric: i is in the list of every j in its list
        for j in lst[rng.integers(0, lst.size, 50)] if lst.size else []:
            lj = idx[off[j]:off[j + 1]]
            k = np.searchsorted(lj, i)
            assert k < lj.size and lj[k] == i
        c2, i2 = orc.radius_pass("dna", dna, r, (int(i), int(i) + 1))
        assert c2[0] == counts[i] and np.array_equal(i2, lst), i
    assert off[n - 1] > 2 ** 31  # the last rows' ids sit past 2^31
    assert np.array_equal(cells.try_radii([r], q)[0], counts[q])
    checksum = int(idx.sum(dtype=np.uint64))
    cells.close()
    del idx
    monkeypatch.setenv("CASPHY_DNA_FORM", "valu")
    cells = hip.Cells(dna, "dna")
    assert cells.radius(r) == total
    c3, i3 = cells.fetch()
    assert np.array_equal(c3, counts) and int(i3.sum(dtype=np.uint64)) == checksum
