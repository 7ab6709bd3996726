"""GPU parity: the HIP library through its C ABI vs the CPU oracle, bit-exact.

Sizes are chosen so the oracle finishes in seconds; full-size properties are
in test_gpu_fullsize.py.
"""
import numpy as np
import pytest

from casphy_amd import synth

pytestmark = pytest.mark.gpu


def rand_bits(rng, n, W, p):
    return synth.pack_bits(rng.random((n, 32 * W)) < p)


# ------------------------------------------------------------------ scores
@pytest.fixture(params=["mfma-auto", "mfma-lds", "mfma-fp4", "mfma-fp4b", "mfma-fp4d", "mfma-fp4w", "mfma-fp4x",
                        "mfma-fp4y", "mfma-fp4p", "mfma-wide", "mfma-wide8", "mfma-reg", "valu", "valu-k16x8",
                        "valu-k32x4", "valu-k16x4"])
def bitpath(request, monkeypatch):
    """Runs a test on the autotuned int8 matrix-core path (every candidate
    variant runs on the first call of a shape class), on the fixed variants
    (LDS-staged bytes with 128 x 128 or 256 x 256 workgroup tiles, fp4
    nibbles, register-resident bits), on the autotuned VALU popcount tiles and
    on each VALU tile shape forced (kKCxCB: KC words staged per step, CB
    columns per thread)."""
    valu = request.param.startswith("valu")
    monkeypatch.setenv("CASPHY_MFMA", "0" if valu else "1")
    if valu:
        if request.param == "valu":
            monkeypatch.delenv("CASPHY_VALU_VARIANT", raising=False)
            monkeypatch.delenv("CASPHY_AUTOTUNE", raising=False)
        else:
            monkeypatch.setenv("CASPHY_VALU_VARIANT", request.param[5:])
        return request.param
    if request.param == "mfma-auto":
        monkeypatch.delenv("CASPHY_MFMA_VARIANT", raising=False)
        monkeypatch.delenv("CASPHY_AUTOTUNE", raising=False)
    else:
        variant = {"mfma-reg": "reg", "mfma-wide": "wide", "mfma-wide8": "wide8", "mfma-fp4": "fp4",
                   "mfma-fp4b": "fp4b", "mfma-fp4d": "fp4d", "mfma-fp4w": "fp4w",
                   "mfma-fp4x": "fp4x", "mfma-fp4y": "fp4y", "mfma-fp4p": "fp4p"}.get(request.param, "64x2")
        monkeypatch.setenv("CASPHY_MFMA_VARIANT", variant)
    return request.param


@pytest.mark.parametrize("n,W", [(1, 1), (5, 1), (127, 2), (128, 4), (129, 3), (300, 5), (515, 32)])
def test_score_bit_matrix(hip, orc, bitpath, n, W):
    rng = np.random.default_rng(n * 7 + W)
    mat = rand_bits(rng, n, W, 0.15)
    S = hip.score_bit_matrix(mat)
    assert np.array_equal(S, orc.score_bit_matrix(mat).astype(np.int32))


@pytest.mark.parametrize("sup", ["0", "3"])
def test_fp4p_many_tiles(hip, orc, monkeypatch, sup):
    """The persistent 256 x 256 fp4 tiles over many more tiles than
    workgroups, so every workgroup runs several tiles and prefetches across
    tile boundaries: 6,100 cells x 640 sites (5 stages, a ragged last row
    block), scores and radius lists against the oracle, in plain and in 3 x 3
    super-tile order."""
    monkeypatch.setenv("CASPHY_MFMA", "1")
    monkeypatch.setenv("CASPHY_MFMA_VARIANT", "fp4p")
    monkeypatch.setenv("CASPHY_FP4P_SUPER", sup)
    rng = np.random.default_rng(61)
    mat = rand_bits(rng, 6100, 20, 0.2)
    S = hip.score_bit_matrix(mat)
    assert np.array_equal(S, orc.score_bit_matrix(mat).astype(np.int32))
    del S
    for r in (30, 45):
        counts, idx = hip.radius_neighbor(np.ascontiguousarray(mat.T), r)
        c2, i2 = orc.radius_bit(mat, r)
        assert np.array_equal(counts, c2) and np.array_equal(idx, i2)


def test_fp4p_other_epilogues(hip, orc, monkeypatch):
    """The persistent 256 x 256 fp4 tiles forced for the epilogues they reach
    only through the autotuner otherwise: transfer minima over a rectangle of
    more tiles than workgroups (6,000 x 5,000 clade rows of 640 tips, 24 x 20
    tiles), tipMerge pair counts (dense form, 512 sites) and the voteRoot2
    argmax (3,000 nodes x 500 samples, 1,000 positions), against the oracle."""
    monkeypatch.setenv("CASPHY_MFMA", "1")
    monkeypatch.setenv("CASPHY_MFMA_VARIANT", "fp4p")
    rng = np.random.default_rng(91)
    ntip = 640
    Y = synth.pack_bits(rng.random((6000, ntip)) < 0.3)
    X = synth.pack_bits(rng.random((5000, ntip)) < rng.random((5000, 1)))
    for g, w in zip(hip.transfer_dis(Y, X, ntip), orc.transfer(Y, X, ntip)):
        assert np.array_equal(g, w)
    monkeypatch.setenv("CASPHY_TIPMERGE", "dense")
    mat = synth.binary_matrix(seed=1500, n=1500, nsite=512, p=0.01)
    counts, idx = hip.radius_neighbor(np.ascontiguousarray(mat.T), 12)
    for p in (0.0, 0.5, 0.8):
        assert np.array_equal(hip.tip_merge(counts, idx, p), orc.tip_merge(counts, idx, p)), p
    st, edge, vmat, vidx = synth.vote_inputs(3000, 3000, 500, 1000)
    root, nsh = hip.vote_root(st, edge, vmat, vidx)
    want_root, want_n = orc.vote_root(st, edge, vmat, vidx)
    assert np.array_equal(nsh, want_n) and np.array_equal(root, want_root)


def test_score_bit_past_fp32_exact_sums(hip, monkeypatch):
    """Rows of 2^24 + 1,024 sites (sums above 2^24, where fp32 sums stop being
    exact): the matrix-core tiles take the int8 form's int32 sums even when
    an fp4 variant is asked for, and the scores equal the exact popcounts
    (odd values above 2^24 included)."""
    monkeypatch.setenv("CASPHY_MFMA", "1")
    monkeypatch.setenv("CASPHY_MFMA_VARIANT", "fp4")
    W = (1 << 19) + 32
    mat = np.full((3, W), 0xFFFFFFFF, np.uint32)
    mat[1, 5] = 0xFFFFFFFE          # one site fewer
    mat[2, :W // 2] = 0x55555555    # half the sites in the first half
    S = hip.score_bit_matrix(mat)
    pop = lambda a: int(np.unpackbits(a.view(np.uint8)).sum())  # noqa: E731
    want = np.array([[pop(mat[i] & mat[j]) for j in range(3)] for i in range(3)], np.int64)
    assert want[0, 1] == (1 << 24) + 1023
    assert np.array_equal(S.astype(np.int64), want)


def test_autotuner_report(hip, monkeypatch):
    monkeypatch.delenv("CASPHY_MFMA_VARIANT", raising=False)
    monkeypatch.setenv("CASPHY_MFMA", "1")
    rng = np.random.default_rng(11)
    mat = rand_bits(rng, 700, 8, 0.1)
    S1 = hip.score_bit_matrix(mat)     # first call of this shape class: tuning
    S2 = hip.score_bit_matrix(mat)     # cached choice
    assert np.array_equal(S1, S2)
    rep = [e for e in hip.mfma_tuning() if e.get("epilogue") == 1 and e["log2_rows_a"] == 10]  # score epilogue
    assert rep and all(e["chosen"] in e["ms"] for e in rep)
    assert all(t > 0 for e in rep for t in e["ms"].values())


def test_score_bit_weighted(hip, orc):
    rng = np.random.default_rng(1)
    mat = rand_bits(rng, 200, 4, 0.2)
    for w in (rng.integers(0, 9, 128), np.full(128, 3), np.zeros(128), rng.integers(0, 1000, 128)):
        S = hip.score_bit_matrix(mat, w.astype(np.float64))
        assert np.array_equal(S, orc.score_bit_matrix(mat, w.astype(np.float32)).astype(np.int32))


def test_weights_outside_contract_rejected(hip):
    mat = np.ones((4, 1), np.uint32)
    for w in (np.full(32, 0.5), np.full(32, -1.0), np.full(32, 1e6)):
        with pytest.raises(hip.CasphyError) as e:
            hip.score_bit_matrix(mat, w)
        assert e.value.code == -4


def test_score_indel_matrix_config1(hip, orc):
    """The indel configuration's full 1,024 x 1,024 score matrix (200 sites,
    alleles from 100 ids) against the oracle's countApoFloat over all pairs."""
    mat = synth.indel_matrix(seed=1)
    w = np.ones((mat.max(), mat.shape[1]), dtype=np.float32)
    S = hip.score_indel_matrix(mat)
    assert S.shape == (1024, 1024)
    assert np.array_equal(S, orc.score_indel_matrix(mat, w).astype(np.int32))


def test_score_indel_matrix(hip, orc):
    mat = synth.indel_matrix(seed=3, depth=7, nsite=60)
    w = np.ones((mat.max(), 60), dtype=np.float32)
    n = mat.shape[0]
    S = hip.score_indel_matrix(mat)
    ref = np.array([[orc.score_indel(mat[i], mat[j], w) for j in range(n)] for i in range(n)], np.int32)
    assert np.array_equal(S, ref)
    rng = np.random.default_rng(4)
    wi = rng.integers(0, 7, w.shape)
    S2 = hip.score_indel_matrix(mat, wi.astype(np.float64))
    i, j = 5, 77
    assert S2[i, j] == orc.score_indel(mat[i], mat[j], wi.astype(np.float32))
    assert np.array_equal(S2, S2.T)


# ------------------------------------------------------------------ radius
@pytest.mark.parametrize("n,W,p,r", [(1, 1, 0.3, 1), (33, 1, 0.2, 2), (250, 4, 0.05, 3), (700, 8, 0.03, 5),
                                     (1000, 2, 0.1, 0), (257, 3, 0.1, 1000), (4999, 2, 0.05, 2)])
def test_radius_bit(hip, orc, bitpath, n, W, p, r):
    rng = np.random.default_rng(n + W)
    mat = rand_bits(rng, n, W, p)
    counts, idx = hip.radius_neighbor(np.ascontiguousarray(mat.T), r)
    c2, i2 = orc.radius_bit(mat, r)
    assert np.array_equal(counts, c2)
    assert np.array_equal(idx, i2)


def test_radius_bit_weighted(hip, orc):
    rng = np.random.default_rng(9)
    mat = rand_bits(rng, 300, 2, 0.1)
    w = rng.integers(0, 5, 64).astype(np.float64)
    counts, idx = hip.radius_neighbor(np.ascontiguousarray(mat.T), 7, w)
    c2, i2 = orc.radius_bit(mat, 7, w.astype(np.float32))
    assert np.array_equal(counts, c2) and np.array_equal(idx, i2)


def test_radius_bit_synthetic_config(hip, orc, bitpath):
    mat = synth.binary_matrix(seed=2, n=1500, nsite=1024, p=0.005)
    for r in (1, 4, 16):
        counts, idx = hip.radius_neighbor(np.ascontiguousarray(mat.T), r)
        c2, i2 = orc.radius_bit(mat, r)
        assert np.array_equal(counts, c2) and np.array_equal(idx, i2)


@pytest.fixture(params=["auto", "mfma", "mfma-fp4p", "valu", "valu-k48x8", "valu-k48x4"])
def dnaform(request, monkeypatch):
    """ACGT data on the one-hot matrix-core form (autotuned tiles, or the
    persistent 256 x 256 tiles forced), on the 2-bit + mask VALU form
    (autotuned, or with the 48-word stage in 128 x 128 or 128 x 64 tiles
    forced), or on the form the resident cells' tuner picks."""
    monkeypatch.setenv("CASPHY_MFMA", "1")
    monkeypatch.delenv("CASPHY_MFMA_VARIANT", raising=False)
    monkeypatch.delenv("CASPHY_VALU_VARIANT", raising=False)
    if request.param == "auto":
        monkeypatch.delenv("CASPHY_DNA_FORM", raising=False)
    elif request.param == "mfma-fp4p":
        monkeypatch.setenv("CASPHY_DNA_FORM", "mfma")
        monkeypatch.setenv("CASPHY_MFMA_VARIANT", "fp4p")
    elif request.param.startswith("valu-"):
        monkeypatch.setenv("CASPHY_DNA_FORM", "valu")
        monkeypatch.setenv("CASPHY_VALU_VARIANT", request.param[5:])
    else:
        monkeypatch.setenv("CASPHY_DNA_FORM", request.param)
    return request.param


@pytest.mark.parametrize("n,ncol", [(2, 3), (100, 90), (600, 1000), (257, 1201)])
def test_radius_dna_2bit(hip, orc, dnaform, n, ncol):
    mat = synth.dna_matrix(seed=n, n=n, ncol=ncol)
    for r in (1, ncol // 2, int(ncol * 0.9)):
        counts, idx = hip.radius_neighbor_dna(mat, r)
        c2, i2 = orc.radius_dna(mat, r)
        assert np.array_equal(counts, c2) and np.array_equal(idx, i2)


def test_radius_dna_iupac_nibble_path(hip, orc):
    rng = np.random.default_rng(10)
    codes = rng.choice(np.array([1, 2, 4, 8, 3, 15, 0x10, 5], np.uint8), (300, 203))
    mat = synth.pack_bytes4(codes)
    for r in (1, 60, 120):
        counts, idx = hip.radius_neighbor_dna(mat, r)
        c2, i2 = orc.radius_dna(mat, r)
        assert np.array_equal(counts, c2) and np.array_equal(idx, i2)


def test_try_radius(hip, orc, bitpath):
    mat = synth.binary_matrix(seed=4, n=900, nsite=512, p=0.01)
    bmat = np.ascontiguousarray(mat.T)
    q = synth.query_rows(5, 900, 300)
    radii = np.arange(1, 17)
    got = hip.try_radii(bmat, radii, q)
    for k, r in enumerate(radii):
        assert np.array_equal(got[k], orc.try_radius_bit(mat, r, q)), r
    assert np.array_equal(hip.try_radius(bmat, 3, q), orc.try_radius_bit(mat, 3, q))


@pytest.mark.parametrize("n", [901, 1024])
def test_try_radii_radius_sets(hip, orc, n):
    # repeated radii, radius 0, the largest radius of the score table (1022),
    # radii beyond it and negative radii (per-radius compares); 901 cells leave
    # score rows off the 16-byte grid
    mat = synth.pack_bits(np.random.default_rng(n).random((n, 3000)) < 0.35)  # pair scores ~367, self ~1050
    bmat = np.ascontiguousarray(mat.T)
    q = synth.query_rows(6, n, 70)
    for radii in ([0, 5, 360, 360, 380, 400], [400, 1022, 1022], [5, 380, 1023], [-1, 2, 360], list(range(340, 420, 5))):
        got = hip.try_radii(bmat, np.array(radii), q)
        for k, r in enumerate(radii):
            assert np.array_equal(got[k], orc.try_radius_bit(mat, r, q)), (radii, r)


def autosplit_rule(counts, radii, bound):
    """The autoSplit stopping rule on host: first radius whose mean count c has
    log10(max(c, 1e-30)) <= bound, else -1."""
    for k, r in enumerate(radii):
        if np.log10(max(counts[k].mean(), 1e-30)) <= bound:
            return int(r)
    return -1


def test_cells_autosplit_bit(hip, orc, bitpath):
    mat = synth.binary_matrix(seed=23, n=900, nsite=256, p=0.01)
    cells = hip.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
    q = synth.query_rows(5, 900, 200)
    radii = np.arange(1, 17) * 2
    ref = np.stack([orc.try_radius_bit(mat, int(r), q) for r in radii])
    seen = set()
    # larger, then smaller radii: the resident list buffer grows between calls
    for bound in (0.5, -40.0, 3.0, 1.5, 0.5):
        want = autosplit_rule(ref, radii, bound)
        r, total, cnt = cells.autosplit(radii, q, bound, counts=True)
        assert np.array_equal(cnt, ref)
        assert r == want, bound
        if want < 0:
            assert total == 0
            continue
        seen.add(want)
        counts, idx = cells.fetch()
        c2, i2 = orc.radius_bit(mat, want)
        assert total == idx.size and np.array_equal(counts, c2) and np.array_equal(idx, i2)
        assert cells.autosplit(radii, q, bound) == (r, total)
    assert len(seen) >= 2
    # repeated identical rounds: the second is captured into a graph, later ones replay it
    want = autosplit_rule(ref, radii, 1.5)
    c2, i2 = orc.radius_bit(mat, want)
    for rep in range(4):
        r, total, cnt = cells.autosplit(radii, q, 1.5, counts=True)
        counts, idx = cells.fetch()
        assert r == want and np.array_equal(cnt, ref), rep
        assert np.array_equal(counts, c2) and np.array_equal(idx, i2), rep


def test_cells_autosplit_dna(hip, orc, dnaform):
    mat = synth.dna_matrix(seed=11, n=400, ncol=300)
    cells = hip.Cells(mat, "dna")
    q = synth.query_rows(12, 400, 130)
    radii = np.array([100, 200, 250, 280, 290, 299])
    ref = np.stack([orc.try_radius_dna(mat, int(r), q) for r in radii])
    for bound in (1.0, 2.0, 2.7):
        want = autosplit_rule(ref, radii, bound)
        r, total, cnt = cells.autosplit(radii, q, bound, counts=True)
        assert np.array_equal(cnt, ref) and r == want
        if want >= 0:
            counts, idx = cells.fetch()
            c2, i2 = orc.radius_dna(mat, want)
            assert np.array_equal(counts, c2) and np.array_equal(idx, i2)


def test_cells_autosplit_limit_exact(hip):
    """The device comparison equals the host rule at the boundary: bounds set to
    log10 of the exact mean count of one radius."""
    mat = synth.binary_matrix(seed=24, n=700, nsite=256, p=0.01)
    cells = hip.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
    q = synth.query_rows(6, 700, 333)
    radii = np.arange(1, 17)
    cnt = cells.try_radii(radii, q)
    for k in (2, 5, 9):
        m = cnt[k].mean()
        if m <= 0:
            continue
        for bound in (np.log10(m), np.nextafter(np.log10(m), -np.inf), np.nextafter(np.log10(m), np.inf)):
            r, _ = cells.autosplit(radii, q, float(bound))
            assert r == autosplit_rule(cnt, radii, float(bound)), (k, bound)


@pytest.mark.parametrize("kind", ["bit", "dna"])
def test_cells_autosplit_none_qualifies(hip, orc, kind):
    """A round in which no radius qualifies skips the mask tiles, counts and
    compaction on the device: empty lists, singleton tipMerge labels, and the
    next qualifying round (replayed from its graph) is exact again."""
    if kind == "bit":
        mat = synth.binary_matrix(seed=31, n=900, nsite=1024, p=0.01)
        cells = hip.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
        radii = np.arange(1, 17)
        ref_lists = lambda r: orc.radius_bit(mat, r)  # noqa: E731
        ref_try = lambda r, q: orc.try_radius_bit(mat, r, q)  # noqa: E731
    else:
        mat = synth.dna_matrix(seed=31, n=900, ncol=500)
        cells = hip.Cells(mat, "dna")
        radii = np.arange(200, 360, 10)
        ref_lists = lambda r: orc.radius_dna(mat, r)  # noqa: E731
        ref_try = lambda r, q: orc.try_radius_dna(mat, r, q)  # noqa: E731
    q = synth.query_rows(32, mat.shape[0], 200)
    ref = np.stack([ref_try(int(r), q) for r in radii])
    good = float(np.log10(ref[len(radii) // 2].mean()))  # some radius qualifies
    want = autosplit_rule(ref, radii, good)
    assert 0 <= want <= radii[len(radii) // 2]
    c2, i2 = ref_lists(want)
    for rep in range(3):
        r, total = cells.autosplit(radii, q, good)
        counts, idx = cells.fetch()
        assert r == want and np.array_equal(counts, c2) and np.array_equal(idx, i2), rep
        r, total = cells.autosplit(radii, q, -40.0)
        assert r == -1 and total == 0, rep
        counts, idx = cells.fetch()
        assert not counts.any() and idx.size == 0, rep
        member = cells.tip_merge(0.8)
        assert np.array_equal(member, np.arange(mat.shape[0])), rep


def test_try_radius_dna(hip, orc, dnaform):
    mat = synth.dna_matrix(seed=11, n=400, ncol=300)
    q = synth.query_rows(12, 400, 130)
    radii = [100, 200, 250, 280, 299]
    got = hip.try_radii_dna(mat, radii, q)
    for k, r in enumerate(radii):
        assert np.array_equal(got[k], orc.try_radius_dna(mat, r, q))


# ------------------------------------------------------------------ joins
@pytest.mark.parametrize("n,W,p", [(2, 1, 0.3), (3, 1, 0.0), (17, 1, 0.2), (130, 2, 0.1), (400, 8, 0.02)])
@pytest.mark.parametrize("anc", [0, 1])
def test_join_bit(hip, orc, bitpath, n, W, p, anc):
    rng = np.random.default_rng(n * 3 + anc)
    mat = rand_bits(rng, n, W, p)
    e, ns, nb = hip.join_bit(mat, None, anc)
    e2, ns2, nb2 = orc.join_bit(mat, ancestral_type=anc)
    assert np.array_equal(e, e2)
    assert np.array_equal(ns, ns2) and np.array_equal(nb, nb2)


@pytest.fixture(params=["fstep", "fstep-onehot", "fused", "two-kernel", "fstep-looped", "fstep-onehot-looped",
                        "fused-looped", "two-kernel-looped"])
def joinstep(request, monkeypatch):
    """Single-tree join in its three step forms: one launch per step that
    repeats the selection in every workgroup (the default, "fstep"; indel rows
    as byte codes, or as the one-hot planes with "onehot"), one launch whose
    last workgroup selects ("fused"), or rescore + select launches
    ("two-kernel"); "looped": 5 workgroups, so each one loops over several
    groups of live slots (the form large joins take)."""
    monkeypatch.setenv("CASPHY_JOIN_CODES", "0" if "onehot" in request.param else "1")
    form = request.param.replace("-looped", "").replace("-onehot", "")
    monkeypatch.setenv("CASPHY_JOIN_FUSED", "1" if form == "fused" else "0")
    if form == "two-kernel":
        monkeypatch.setenv("CASPHY_JOIN_STEP", "two")
    else:
        monkeypatch.delenv("CASPHY_JOIN_STEP", raising=False)
    if request.param.endswith("looped"):
        monkeypatch.setenv("CASPHY_JOIN_RESCORE_WG", "5")
        monkeypatch.setenv("CASPHY_JOIN_WG", "5")
    else:
        monkeypatch.delenv("CASPHY_JOIN_RESCORE_WG", raising=False)
        monkeypatch.delenv("CASPHY_JOIN_WG", raising=False)
    return request.param


def test_join_bit_synthetic(hip, orc, joinstep):
    mat = synth.binary_matrix(seed=21, n=1200, nsite=512, p=0.01)
    e, ns, _ = hip.join_bit(mat)
    e2, ns2, _ = orc.join_bit(mat, lazy=True)
    assert np.array_equal(e, e2) and np.array_equal(ns, ns2)


def test_join_bit_weighted(hip, orc):
    rng = np.random.default_rng(22)
    mat = rand_bits(rng, 150, 2, 0.1)
    w = rng.integers(0, 6, 64).astype(np.float64)
    e, _, _ = hip.join_bit(mat, w, 1)
    e2, _, _ = orc.join_bit(mat, w.astype(np.float32), 1)
    assert np.array_equal(e, e2)


@pytest.mark.parametrize("min_size,ratio", [(0, 2.0), (4, 1.5)])
def test_join_indel(hip, orc, bitpath, min_size, ratio):
    mat = synth.indel_matrix(seed=31, depth=8, nsite=120)
    n = mat.shape[0]
    w = np.ones((mat.max(), 120))
    ts = np.random.default_rng(5).integers(1, 6, n).astype(np.uint32)
    e, ns, br = hip.join_indel(mat, w, ts, min_size, ratio)
    e2, ns2, br2 = orc.join_indel(mat, w.astype(np.float32), ts, min_size, ratio)
    assert np.array_equal(e, e2) and np.array_equal(ns, ns2) and np.array_equal(br, br2)


@pytest.mark.parametrize("kmax,ncol", [(3, 37), (254, 41), (300, 33)])
def test_join_indel_codes_edges(hip, orc, kmax, ncol):
    """Indel joins on byte codes: many missing states, alleles up to 254 (the
    largest a code holds; 300 falls back to the one-hot planes), site counts
    that are not a multiple of 4; unit and all-one weights (what R passes),
    and node states decoded from the codes, against the oracle."""
    rng = np.random.default_rng(kmax + ncol)
    n = 300
    mat = rng.integers(1, kmax + 1, (n, ncol)).astype(np.int32)
    mat[rng.random((n, ncol)) < 0.45] = 0
    mat[rng.random((n, ncol)) < 0.3] = -1
    mat[:, 0] = kmax  # the largest allele occurs
    w = np.ones((kmax, ncol), np.float32)
    for weight in (None, w.astype(np.float64)):
        e, ns, br = hip.join_indel(mat, weight)
        e2, ns2, br2 = orc.join_indel(mat, w)
        assert np.array_equal(e, e2) and np.array_equal(ns, ns2) and np.array_equal(br, br2)
    # the same rows as three jobs of one persistent launch, with the size rule
    parts = [mat[:120], mat[120:121 + 60], mat[181:]]
    ts = [np.random.default_rng(k).integers(1, 6, p.shape[0]).astype(np.uint32) for k, p in enumerate(parts)]
    batch = hip.join_indel_batch(parts, w.astype(np.float64), ts, 3, 1.5)
    for p, t, e in zip(parts, ts, batch):
        e2, _, _ = orc.join_indel(p, w, t, 3, 1.5)
        assert np.array_equal(e, e2)


def test_join_indel_config1(hip, orc, joinstep):
    mat = synth.indel_matrix(seed=1)
    e, ns, br = hip.join_indel(mat)
    w = np.ones((mat.max(), mat.shape[1]), np.float32)
    e2, ns2, br2 = orc.join_indel(mat, w)
    assert np.array_equal(e, e2) and np.array_equal(ns, ns2) and np.array_equal(br, br2)


def test_cladewise(hip, orc):
    mat = synth.binary_matrix(seed=7, n=500, nsite=128, p=0.02)
    e, _, _ = orc.join_bit(mat)
    eo, io = hip.cladewise(e, 500)
    eo2, io2 = orc.cladewise(e, 500)
    assert np.array_equal(eo, eo2) and np.array_equal(io, io2)


def test_join_batch_equals_single(hip):
    mats = [synth.binary_matrix(seed=40 + k, n=int(nk), nsite=256, p=0.02) for k, nk in enumerate([2, 57, 300, 129])]
    batch = hip.join_bit_batch(mats)
    for m, e in zip(mats, batch):
        e1, _, _ = hip.join_bit(m)
        assert np.array_equal(e, e1)


def test_join_batch_weighted_and_types(hip):
    rng = np.random.default_rng(41)
    mats = [rand_bits(rng, int(k), 3, 0.1) for k in (5, 64, 200)]
    w = rng.integers(0, 4, 96).astype(np.float64)
    for anc in (0, 1):
        batch = hip.join_bit_batch(mats, w, anc)
        for m, e in zip(mats, batch):
            e1, _, _ = hip.join_bit(m, w, anc)
            assert np.array_equal(e, e1)


@pytest.mark.parametrize("rows", ["codes", "onehot"])
def test_join_indel_batch_bootstrap(hip, orc, monkeypatch, rows):
    """Bootstrap replicates of the indel configuration (sites resampled with
    replacement) in one persistent launch, each equal to its own oracle join;
    rows as byte codes (the default) or as the one-hot planes."""
    monkeypatch.setenv("CASPHY_JOIN_CODES", "1" if rows == "codes" else "0")
    base = synth.indel_matrix(seed=1, depth=8)
    rng = np.random.default_rng(42)
    reps = [base[:, rng.integers(0, base.shape[1], base.shape[1])] for _ in range(6)]
    reps.append(base[:3])
    ts = [np.ones(r.shape[0], np.uint32) for r in reps]
    batch = hip.join_indel_batch(reps, None, ts, 0, 2.0)
    for r, e in zip(reps, batch):
        w = np.ones((max(1, r.max()), r.shape[1]), np.float32)
        e2, _, _ = orc.join_indel(r, w)
        assert np.array_equal(e, e2)


def test_join_indel_bootstrap(hip, orc):
    """Column-resample replicates built on the device equal the batched joins
    of the same replicates built on the host, and the oracle's joins of two of
    them: byte codes (the simulator's 200 sites) and one-hot planes (a site of
    300 alleles), replicates of 256 and 37 cells."""
    rng = np.random.default_rng(44)
    for base in (synth.indel_matrix(seed=5, depth=8), synth.indel_matrix(seed=6, depth=8)[:37, :60].copy()):
        if base.shape[0] == 37:
            base[:, 7] = rng.integers(-1, 301, base.shape[0])  # more alleles than a byte code holds
        n, ncol = base.shape
        cols = rng.integers(0, ncol, (5, ncol)).astype(np.uint32)
        cols[0] = np.arange(ncol)
        got = hip.join_indel_bootstrap(base, cols)
        reps = [np.ascontiguousarray(base[:, c]) for c in cols]
        want = hip.join_indel_batch(reps)
        assert all(np.array_equal(g, w) for g, w in zip(got, want))
        for r in (0, 3):
            w = np.ones((max(1, int(reps[r].max())), ncol), np.float32)
            e2, _, _ = orc.join_indel(reps[r], w)
            assert np.array_equal(got[r], e2)


# ------------------------------------------------------------------ transfer
@pytest.mark.parametrize("ntip,ny,nx", [(64, 62, 62), (75, 20, 300), (1000, 130, 129), (32, 1, 1)])
def test_transfer(hip, orc, bitpath, ntip, ny, nx):
    rng = np.random.default_rng(ntip + ny)
    W = (ntip + 31) // 32

    def bits(m):
        b = rng.random((m, ntip)) < rng.random()
        return synth.pack_bits(np.pad(b, ((0, 0), (0, 32 * W - ntip))))
    Y, X = bits(ny), bits(nx)
    got = hip.transfer_dis(Y, X, ntip)
    want = orc.transfer(Y, X, ntip)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_transfer_trees(hip, orc, bitpath):
    rng = np.random.default_rng(3)
    parent, leaves = synth.yule_parent(700, rng)
    ref = synth.clade_bitsets(parent, leaves, 700)
    boot = synth.clade_bitsets(synth.spr(parent, leaves, 20, rng), leaves, 700)
    got = hip.transfer_dis(ref, boot, 700)
    want = orc.transfer(ref, boot, 700)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_transfer_batch(hip, orc, bitpath):
    rng = np.random.default_rng(77)
    ntip = 500
    parent, leaves = synth.yule_parent(ntip, rng)
    ref = synth.clade_bitsets(parent, leaves, ntip)
    boots = [synth.clade_bitsets(synth.spr(parent, leaves, k, rng), leaves, ntip) for k in (1, 10, 40)]
    boots.append(boots[0][:7])
    ymin, ysize, xm, xs = hip.transfer_dis_batch(ref, boots, ntip)
    for b, X in enumerate(boots):
        y2, ys2, x2, xs2 = orc.transfer(ref, X, ntip)
        assert np.array_equal(ymin[b], y2) and np.array_equal(ysize, ys2)
        assert np.array_equal(xm[b], x2) and np.array_equal(xs[b], xs2)
    # the same replicates back to back in one array
    y3, ys3, xm3, xs3 = hip.transfer_dis_batch(ref, np.concatenate(boots), ntip, [b.shape[0] for b in boots])
    assert np.array_equal(y3, ymin) and np.array_equal(ys3, ysize)
    assert all(np.array_equal(a, b) for a, b in zip(xm3, xm)) and all(np.array_equal(a, b) for a, b in zip(xs3, xs))


def test_transfer_batch_streamed_groups(hip, orc, monkeypatch):
    """Replicates streamed through the device one group at a time (a budget
    that holds one replicate per group, empty replicates included) give the
    oracle's minima, for bitsets in and for trees in on the bitset tiles."""
    monkeypatch.setenv("CASPHY_TBE_BUDGET_MB", "1")
    rng = np.random.default_rng(78)
    ntip = 3000
    parent, leaves = synth.yule_parent(ntip, rng)
    ref = synth.clade_bitsets(parent, leaves, ntip)
    boots = [synth.clade_bitsets(synth.spr(parent, leaves, k, rng), leaves, ntip) for k in (3, 30)]
    boots.insert(1, boots[0][:0])
    ymin, ysize, xm, xs = hip.transfer_dis_batch(ref, boots, ntip)
    for b, X in enumerate(boots):
        y2, ys2, x2, xs2 = orc.transfer(ref, X, ntip)
        assert np.array_equal(ymin[b], y2) and np.array_equal(ysize, ys2), b
        assert np.array_equal(xm[b], x2) and np.array_equal(xs[b], xs2), b
    monkeypatch.setenv("CASPHY_TBE", "bitsets")
    ref_e = synth.ape_edges(parent, leaves)[:2]
    reps = [synth.ape_edges(synth.spr(parent, leaves, 50, rng), leaves)[:2] for _ in range(3)]
    reps.insert(1, synth.ape_edges(np.array([-1] + [0] * ntip), np.arange(1, ntip + 1))[:2])  # no rows
    y3, ys3, x3, xs3 = hip.transfer_trees(ref_e[0], ref_e[1], [e for e, _ in reps], [n for _, n in reps], ntip)
    monkeypatch.delenv("CASPHY_TBE_BUDGET_MB")
    monkeypatch.delenv("CASPHY_TBE")
    y4, ys4, x4, xs4 = hip.transfer_trees(ref_e[0], ref_e[1], [e for e, _ in reps], [n for _, n in reps], ntip)
    assert np.array_equal(y3, y4) and np.array_equal(ys3, ys4)
    assert all(np.array_equal(a, b) for a, b in zip(x3, x4)) and all(np.array_equal(a, b) for a, b in zip(xs3, xs4))


# ------------------------------------------------------------------ scan form
@pytest.mark.parametrize("n,W", [(700, 256), (333, 300), (1025, 1024)])
def test_try_radius_scan_form(hip, orc, monkeypatch, n, W):
    """Up to four queries over wide rows take one HBM pass (scan form); its
    counts equal the oracle and the tile path."""
    rng = np.random.default_rng(n + W)
    mat = rand_bits(rng, n, W, 0.02)
    bmat = np.ascontiguousarray(mat.T)
    radii = np.arange(1, 17) * max(1, W // 64)
    for R in (1, 3, 4):
        q = synth.query_rows(R, n, R)
        got = hip.try_radii(bmat, radii, q)
        for k, r in enumerate(radii):
            assert np.array_equal(got[k], orc.try_radius_bit(mat, int(r), q)), (R, r)
        monkeypatch.setenv("CASPHY_SCAN", "0")
        assert np.array_equal(hip.try_radii(bmat, radii, q), got)
        monkeypatch.delenv("CASPHY_SCAN")


def test_radius_full_rows(hip, orc, bitpath):
    """Rows whose mask chunks are all ones (radius 0 and 1 on dense rows) take
    the run-of-ids compaction path, also when a resident buffer has to grow."""
    rng = np.random.default_rng(77)
    mat = rand_bits(rng, 2100, 1, 0.6)
    bmat = np.ascontiguousarray(mat.T)
    for r in (0, 1, 9):
        counts, idx = hip.radius_neighbor(bmat, r)
        c2, i2 = orc.radius_bit(mat, r)
        assert np.array_equal(counts, c2) and np.array_equal(idx, i2), r
    cells = hip.Cells(bmat, "bit", wordmajor=True)
    q = synth.query_rows(3, 2100, 64)
    for radii in ([12, 16], [0, 1]):  # small lists first, then every pair
        r, total = cells.autosplit(np.array(radii), q, 10.0)
        counts, idx = cells.fetch()
        c2, i2 = orc.radius_bit(mat, r)
        assert r == radii[0] and np.array_equal(counts, c2) and np.array_equal(idx, i2)


def test_rcpp_radius_shims_one_pass(hip, orc, monkeypatch):
    """radiusWeightedParallel / radiusWeightedDNAParallel, called once per row
    as R calls them: one device pass per matrix and radius, every row's
    count and bitmask equal to the oracle's list."""
    from casphy_amd import api
    calls = []
    real_bit, real_dna = api.radius_neighbor, api.radius_neighbor_dna
    monkeypatch.setattr(api, "radius_neighbor", lambda *a, **k: calls.append("bit") or real_bit(*a, **k))
    monkeypatch.setattr(api, "radius_neighbor_dna", lambda *a, **k: calls.append("dna") or real_dna(*a, **k))
    mat = synth.binary_matrix(seed=61, n=300, nsite=512, p=0.01)
    c2, i2 = orc.radius_bit(mat, 3)
    off = np.concatenate([[0], np.cumsum(c2, dtype=np.int64)])
    bmat = np.ascontiguousarray(mat.T)
    for i in range(300):
        out = api.radiusWeightedParallel(i, bmat, 1, None, 3.0)
        lst = i2[off[i]:off[i + 1]]
        assert out[0] == lst.size
        bits = np.unpackbits(out[1:].view(np.uint8), bitorder="little")
        assert np.array_equal(np.nonzero(bits)[0], lst), i
    dna = synth.dna_matrix(seed=62, n=200, ncol=300)
    c3, i3 = orc.radius_dna(dna, 200)
    off = np.concatenate([[0], np.cumsum(c3, dtype=np.int64)])
    for i in range(200):
        out = api.radiusWeightedDNAParallel(i, dna, None, 200.0)
        bits = np.unpackbits(out[1:].view(np.uint8), bitorder="little")
        assert out[0] == off[i + 1] - off[i] and np.array_equal(np.nonzero(bits)[0], i3[off[i]:off[i + 1]]), i
    assert calls == ["bit", "dna"]


@pytest.mark.parametrize("kind", ["bit", "dna"])
def test_cells_autosplit_64_radii(hip, orc, kind):
    """Up to 64 radii in one device round (counting passes of 16 radii over
    one score matrix): counts of every radius equal the oracle's, the choice
    equals the host rule's, and the lists the oracle's at that radius; 40
    radii through the plain tryRadius entry point as well."""
    if kind == "bit":
        mat = synth.binary_matrix(seed=71, n=700, nsite=1024, p=0.01)
        cells = hip.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
        radii = np.arange(1, 65)
        ref_try = lambda r, q: orc.try_radius_bit(mat, r, q)  # noqa: E731
        ref_lists = lambda r: orc.radius_bit(mat, r)  # noqa: E731
        plain = lambda rr, q: hip.try_radii(np.ascontiguousarray(mat.T), rr, q)  # noqa: E731
    else:
        mat = synth.dna_matrix(seed=71, n=700, ncol=500)
        cells = hip.Cells(mat, "dna")
        radii = np.arange(100, 420, 5)
        ref_try = lambda r, q: orc.try_radius_dna(mat, r, q)  # noqa: E731
        ref_lists = lambda r: orc.radius_dna(mat, r)  # noqa: E731
        plain = lambda rr, q: hip.try_radii_dna(mat, rr, q)  # noqa: E731
    assert radii.size == 64
    q = synth.query_rows(72, mat.shape[0], 150)
    ref = np.stack([ref_try(int(r), q) for r in radii])
    bound = float(np.log10(max(ref[45].mean(), 1.0)))
    want = autosplit_rule(ref, radii, bound)
    for rep in range(3):
        r, total, cnt = cells.autosplit(radii, q, bound, counts=True)
        assert np.array_equal(cnt, ref) and r == want, rep
        if want >= 0:
            counts, idx = cells.fetch()
            c2, i2 = ref_lists(want)
            assert np.array_equal(counts, c2) and np.array_equal(idx, i2), rep
    assert np.array_equal(cells.try_radii(radii, q), ref)
    assert np.array_equal(plain(radii[:40], q), ref[:40])
