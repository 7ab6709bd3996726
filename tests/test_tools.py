"""The drop-in executables: same argv and file contracts as the reference's
cu/*.cu programs.  GPU tests run them end to end on files and compare every
output byte with the oracle; the CPU tests cover their input error paths."""
import os
import subprocess

import numpy as np
import pytest

from casphy_amd import formats, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "casphy_amd", "bin")


def run(tool, *args):
    return subprocess.run([os.path.join(BIN, tool), *map(str, args)], capture_output=True, text=True, timeout=600)


def have_tools():
    return os.path.exists(os.path.join(BIN, "radiusNeighbor"))


@pytest.mark.skipif(not have_tools(), reason="executables not built (make -C casphy_amd tools)")
@pytest.mark.parametrize("tool,args", [("radiusNeighbor", ["PFX", 3]), ("radiusNeighborDNA", ["PFX", 3]),
                                       ("tryRadius", ["PFX", 3, 10]), ("tryRadiusDNA", ["PFX", 3, 10]),
                                       ("joinEdgeBitReverse", ["PFX", "PFX.w", 1]),
                                       ("joinEdgeIndel", ["PFX", "PFX.w", 2, 0, 2]),
                                       ("tipMerge", ["PFX", 0.8]), ("nodeKeepBit", ["PFX", 1]),
                                       ("nodeKeepIndel", ["PFX", 1]), ("voteRoot2", ["PFX.bin", "PFX"]),
                                       ("tbeSupport", ["PFX.nwk", "PFX.trees", "PFX"])])
def test_missing_input_fails_loudly(tmp_path, tool, args):
    pfx = str(tmp_path / "none")
    r = run(tool, *[a.replace("PFX", pfx) if isinstance(a, str) else a for a in args])
    assert r.returncode != 0
    assert "cannot open" in r.stderr


@pytest.mark.skipif(not have_tools(), reason="executables not built")
def test_truncated_input_fails_loudly(tmp_path):
    pfx = str(tmp_path / "t")
    formats.write_bin(pfx + ".bmat", np.ones((2, 50), np.uint32))
    with open(pfx + ".bmat", "r+b") as f:
        f.truncate(100)
    formats.write_weights(pfx + ".w", np.ones(64))
    r = run("radiusNeighbor", pfx, 1)
    assert r.returncode != 0 and "truncated" in r.stderr


@pytest.mark.gpu
def test_radius_neighbor_tool(tmp_path, orc):
    mat = synth.binary_matrix(seed=50, n=800, nsite=256, p=0.01)
    pfx = str(tmp_path / "b")
    formats.write_bin(pfx + ".bmat", mat.T)          # write.mat2bin of the n x W R matrix
    formats.write_weights(pfx + ".w", np.ones(256))
    r = run("radiusNeighbor", pfx, 2)
    assert r.returncode == 0, r.stderr
    c2, i2 = orc.radius_bit(mat, 2)
    nn = formats.read_bin(pfx + ".nNeighbor")
    assert nn.shape == (1, 800) and np.array_equal(nn[0], c2)
    assert np.array_equal(formats.read_idx(pfx + ".idx"), i2)
    # tryRadius on the same files
    r = run("tryRadius", pfx, 4, 100)
    assert r.returncode == 0, r.stderr
    out = formats.read_bin(pfx + ".nNeighbor")
    assert out.shape == (2, 100)
    assert np.array_equal(out[1], orc.try_radius_bit(mat, 4, out[0]))
    # integer arguments read as atoi reads them: "2.9" and "2x" are 2
    for arg in ("2.9", " 2x"):
        assert run("radiusNeighbor", pfx, arg).returncode == 0
        nn = formats.read_bin(pfx + ".nNeighbor")
        assert np.array_equal(nn[0], c2) and np.array_equal(formats.read_idx(pfx + ".idx"), i2), arg


@pytest.mark.gpu
def test_radius_neighbor_dna_tool(tmp_path, orc):
    mat = synth.dna_matrix(seed=51, n=500, ncol=333)
    pfx = str(tmp_path / "d")
    formats.write_bin(pfx + ".tbmat", mat)
    r = run("radiusNeighborDNA", pfx, 250)
    assert r.returncode == 0, r.stderr
    c2, i2 = orc.radius_dna(mat, 250)
    assert np.array_equal(formats.read_bin(pfx + ".nNeighbor")[0], c2)
    assert np.array_equal(formats.read_idx(pfx + ".idx"), i2)
    r = run("tryRadiusDNA", pfx, 300, 64)
    assert r.returncode == 0, r.stderr
    out = formats.read_bin(pfx + ".nNeighbor")
    assert np.array_equal(out[1], orc.try_radius_dna(mat, 300, out[0]))


@pytest.mark.gpu
@pytest.mark.parametrize("anc", [0, 1])
def test_join_edge_bit_reverse_tool(tmp_path, orc, anc):
    mat = synth.binary_matrix(seed=52, n=400, nsite=320, p=0.02)
    pfx = str(tmp_path / "j")
    formats.write_bin(pfx + ".tbmat", mat)
    formats.write_weights(pfx + ".w", np.ones(320))
    r = run("joinEdgeBitReverse", pfx, pfx + ".w", anc)
    assert r.returncode == 0, r.stderr
    e, ns, _ = orc.join_bit(mat, ancestral_type=anc)
    eo, order = orc.cladewise(e, 400)
    edge2 = formats.read_bin(pfx + ".edge2")
    st2 = formats.read_bin(pfx + ".nodeState2")
    assert edge2.shape == (798, 2) and np.array_equal(edge2, eo)
    assert st2.shape == (399, 10) and np.array_equal(st2, ns[order])


@pytest.mark.gpu
def test_join_edge_indel_tool(tmp_path, orc):
    mat = synth.indel_matrix(seed=53, depth=8, nsite=90)
    n, ncol = mat.shape
    nstate = int(mat.max())
    pfx = str(tmp_path / "i")
    formats.write_bin(pfx + ".tbmat", (mat + 1).astype(np.uint32))
    formats.write_bin(pfx + ".tipSize", np.ones((1, n), np.uint32))
    formats.write_weights(pfx + ".w", np.ones(nstate * ncol))
    r = run("joinEdgeIndel", pfx, pfx + ".w", nstate, 0, 2)
    assert r.returncode == 0, r.stderr
    e, ns, _ = orc.join_indel(mat, np.ones((nstate, ncol), np.float32))
    eo, order = orc.cladewise(e, n)
    assert np.array_equal(formats.read_bin(pfx + ".edge2"), eo)
    assert np.array_equal(formats.read_bin(pfx + ".nodeState2"), (ns[order] + 1).astype(np.uint32))


@pytest.mark.gpu
def test_transfer_dis_stream_tool(tmp_path, orc):
    rng = np.random.default_rng(54)
    parent, leaves = synth.yule_parent(300, rng)
    Y = synth.clade_bitsets(parent, leaves, 300)
    X = synth.clade_bitsets(synth.spr(parent, leaves, 15, rng), leaves, 300)
    yf, xf = str(tmp_path / "y.bin"), str(tmp_path / "x.bin")
    formats.write_bin(yf, Y)
    formats.write_bin(xf, X)
    r = run("TransferDisStream", yf, xf, str(tmp_path / "o"), 300)
    assert r.returncode == 0, r.stderr
    ymin, ysize, xmin, xsize = orc.transfer(Y, X, 300)
    yo = formats.read_bin(str(tmp_path / "o_ymin.bin"))
    xo = formats.read_bin(str(tmp_path / "o_xmin.bin"))
    assert np.array_equal(yo, np.stack([ymin, ysize])) and np.array_equal(xo, np.stack([xmin, xsize]))


@pytest.mark.gpu
def test_single_tip_join_matches_reference_output(tmp_path):
    pfx = str(tmp_path / "one")
    formats.write_bin(pfx + ".tbmat", np.array([[5, 9]], np.uint32))
    formats.write_weights(pfx + ".w", np.ones(64))
    r = run("joinEdgeBitReverse", pfx, pfx + ".w", 1)
    assert r.returncode == 0
    raw = np.fromfile(pfx + ".edge2", "<u4")
    assert list(raw) == [0, 2]
    raw = np.fromfile(pfx + ".nodeState2", "<u4")
    assert list(raw) == [0, 2, 5, 9]


@pytest.mark.gpu
def test_tip_merge_tool(tmp_path, orc):
    """radiusNeighbor then tipMerge on files, as R/CasPhy_functions.R:2024-2026
    runs them; .member holds one label per line."""
    mat = synth.binary_matrix(seed=52, n=900, nsite=512, p=0.01)
    pfx = str(tmp_path / "m")
    formats.write_bin(pfx + ".bmat", mat.T)
    formats.write_weights(pfx + ".w", np.ones(512))
    assert run("radiusNeighbor", pfx, 2).returncode == 0
    r = run("tipMerge", pfx, 0.8)
    assert r.returncode == 0, r.stderr
    assert "tipMerge:" in r.stdout
    member = np.loadtxt(pfx + ".member", dtype=np.int64)
    c2, i2 = orc.radius_bit(mat, 2)
    assert np.array_equal(member, orc.tip_merge(c2, i2, 0.8))


@pytest.mark.gpu
def test_node_keep_tool(tmp_path, orc):
    """joinEdgeBitReverse then nodeKeepBit on the same prefix, as
    R/CasPhy_functions.R:2204,2275-2286 runs them; .retainNode holds the kept
    node ids one per line."""
    mat = synth.binary_matrix(seed=53, n=400, nsite=256, p=0.02)
    pfx = str(tmp_path / "k")
    formats.write_bin(pfx + ".tbmat", mat)
    formats.write_weights(pfx + ".w", np.ones(256))
    assert run("joinEdgeBitReverse", pfx, pfx + ".w", 1).returncode == 0
    out = mat[np.random.default_rng(1).choice(400, 150, replace=False)]
    formats.write_bin(pfx + ".tboutmat", out)
    r = run("nodeKeepBit", pfx, 2)
    assert r.returncode == 0, r.stderr
    eo = formats.read_bin(pfx + ".edge2")
    ns2 = formats.read_bin(pfx + ".nodeState2")
    kept = np.loadtxt(pfx + ".retainNode", dtype=np.int64, ndmin=1) if open(pfx + ".retainNode").read() else []
    assert np.array_equal(np.asarray(kept), orc.node_keep(eo, ns2, out, 2))


@pytest.mark.skipif(not have_tools(), reason="executables not built")
def test_fasta2bin_input_errors(tmp_path):
    """fasta2bin's own messages and exit code 1 (cu/fasta2bin.cu:143-144,178-186,
    258-263,356-359); the text is read before any device work."""
    r = run("fasta2bin", tmp_path / "x.fa")
    assert r.returncode == 1 and "Usage:" in r.stdout
    r = run("fasta2bin", tmp_path / "none.fa", tmp_path / "o")
    assert r.returncode == 1 and "Error: Cannot open file" in r.stderr
    (tmp_path / "e.fa").write_bytes(b">a\n>b\n\n")
    r = run("fasta2bin", tmp_path / "e.fa", tmp_path / "o")
    assert r.returncode == 1 and "No valid sequences found in file" in r.stderr
    (tmp_path / "u.fa").write_bytes(b">a\nACGT\n>b\nACG\n")
    r = run("fasta2bin", tmp_path / "u.fa", tmp_path / "o")
    assert r.returncode == 1 and "same length" in r.stderr
    assert not os.path.exists(str(tmp_path / "o.bin"))
