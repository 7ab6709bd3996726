"""Transfer bootstrap support from Newick text on the GPU: the library's
trees-in path (reader, device clade bitsets, batched transfer, TBE) against
the oracle's reader, clade bitsets, transfer minima and TBE formula, and the
tbeSupport executable end to end."""
import os
import subprocess

import numpy as np
import pytest

from casphy_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "casphy_amd", "bin")


def bootstrap_texts(ntip, moves, seed):
    """A Yule reference and SPR-perturbed replicates as Newick (with lengths
    and internal labels); the replicates list their tips in other orders, so
    they are read back by label."""
    rng = np.random.default_rng(seed)
    parent, leaves = synth.yule_parent(ntip, rng)
    ref = synth.newick(parent, leaves, rng=rng)
    reps = [synth.newick(synth.spr(parent, leaves, k, rng), leaves, rng=rng) for k in moves]
    return ref, "\n".join(reps) + "\n"


def oracle_support(orc, ref_text, rep_text):
    (ye, yn, nt, labels), = orc.read_newick(ref_text)
    reps = orc.read_newick(rep_text, tip_labels=labels)
    Y, ys = orc.clade_bitsets(ye, yn, nt)
    ymin = []
    for e, n, _, _ in reps:
        X, _ = orc.clade_bitsets(e, n, nt)
        y, ysz, _, _ = orc.transfer(Y, X, nt)
        ymin.append(y)
    ymin = np.stack(ymin)
    return orc.tbe_support(ymin, ysz, nt), ysz, ymin


@pytest.mark.parametrize("ntip,moves", [(40, (0, 1, 3)), (700, (2, 20, 70, 200)), (3000, (5, 60))])
def test_transfer_support_from_newick(hip, orc, ntip, moves):
    ref_text, rep_text = bootstrap_texts(ntip, moves, ntip)
    ref = hip.read_newick(ref_text)[0]
    reps = hip.read_newick(rep_text, tip_labels=ref[3])
    sup, ysize, ymin = hip.transfer_support(ref, reps)
    s2, ys2, ym2 = oracle_support(orc, ref_text, rep_text)
    assert np.array_equal(ysize, ys2) and np.array_equal(ymin, ym2)
    assert np.array_equal(np.isnan(sup), np.isnan(s2)) and np.array_equal(sup[~np.isnan(sup)], s2[~np.isnan(s2)])
    if moves[0] == 0:  # an unperturbed replicate finds every split
        assert np.all(ymin[0] == 0)


def test_tbe_support_tool(hip, orc, tmp_path):
    ref_text, rep_text = bootstrap_texts(900, (3, 30, 90), 9)
    (tmp_path / "ref.nwk").write_text(ref_text)
    (tmp_path / "boot.trees").write_text(rep_text)
    pfx = str(tmp_path / "out")
    r = subprocess.run([os.path.join(BIN, "tbeSupport"), str(tmp_path / "ref.nwk"), str(tmp_path / "boot.trees"),
                        pfx], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    lines = open(pfx + "_tbe.txt").read().splitlines()
    assert lines[0] == "node,size,support"
    s2, ys2, _ = oracle_support(orc, ref_text, rep_text)
    (ye, _, nt, _), = orc.read_newick(ref_text)
    nodes = [int(d) for a, d in ye[::-1] if d > nt]  # clade rows: internal children, backward edge order
    assert len(lines) - 1 == len(s2)
    for k, line in enumerate(lines[1:]):
        node, size, sup = line.split(",")
        assert int(node) == nodes[k] and int(size) == ys2[k]
        assert (sup == "NA") if np.isnan(s2[k]) else float(sup) == s2[k]


def brute_trees(orc, ntip, ref, reps):
    """The oracle's transfer minima over the trees' clade bitsets."""
    Y, _ = orc.clade_bitsets(ref[0], ref[1], ntip)
    out = []
    for e, n in reps:
        X, _ = orc.clade_bitsets(e, n, ntip)
        out.append(orc.transfer(Y, X, ntip))
    return out


@pytest.mark.parametrize("form", ["walk", "bitsets"])
def test_transfer_trees_shapes(hip, orc, monkeypatch, form):
    """Trees in, transfer minima out, by the tree walks and by the device clade
    bitsets on the pair tiles: binary SPR replicates, multifurcating trees with
    one-child nodes, a caterpillar and a star, against the oracle."""
    if form == "bitsets":
        monkeypatch.setenv("CASPHY_TBE", "bitsets")
    else:
        monkeypatch.delenv("CASPHY_TBE", raising=False)
    rng = np.random.default_rng(21)
    for ntip in (5, 130, 1100):
        parent, leaves = synth.yule_parent(ntip, rng)
        ref = synth.ape_edges(parent, leaves)[:2]
        reps = [synth.ape_edges(synth.spr(parent, leaves, max(1, ntip // 20), rng), leaves)[:2] for _ in range(2)]
        reps.append(synth.random_tree(rng, ntip, max(1, ntip // 3)))
        reps.append(synth.caterpillar(ntip))
        reps.append(synth.ape_edges(np.array([-1] + [0] * ntip), np.arange(1, ntip + 1))[:2])  # star: no rows
        for r in (ref, synth.random_tree(rng, ntip, max(1, ntip // 2))):
            ymin, ysize, xmin, xsize = hip.transfer_trees(r[0], r[1], [e for e, _ in reps], [n for _, n in reps], ntip)
            for b, (y2, ys2, x2, xs2) in enumerate(brute_trees(orc, ntip, r, reps)):
                assert np.array_equal(ymin[b], y2) and np.array_equal(ysize, ys2), (ntip, b)
                assert np.array_equal(xmin[b], x2) and np.array_equal(xsize[b], xs2), (ntip, b)


def test_transfer_trees_streamed_chunks(hip, monkeypatch):
    """Replicates streamed in chunks of 3 (with host planning overlapping the
    walks) give what one chunk gives."""
    rng = np.random.default_rng(22)
    ntip = 2000
    parent, leaves = synth.yule_parent(ntip, rng)
    ref = synth.ape_edges(parent, leaves)[:2]
    reps = [synth.ape_edges(synth.spr(parent, leaves, 100, rng), leaves)[:2] for _ in range(8)]
    args = (ref[0], ref[1], [e for e, _ in reps], [n for _, n in reps], ntip)
    monkeypatch.setenv("CASPHY_TBE_CHUNK", "3")
    a = hip.transfer_trees(*args)
    monkeypatch.setenv("CASPHY_TBE_CHUNK", "1000")
    b = hip.transfer_trees(*args)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert all(np.array_equal(x, y) for x, y in zip(a[2], b[2]))


def test_transfer_trees_65536_tips(hip, orc, monkeypatch):
    """The batched transfer configuration's tree size: a 65,536-tip Yule tree
    against two 5%-SPR replicates.  The walks equal the bitset pair tiles on
    every row (two independent kernels), and sampled rows equal the oracle's
    brute force."""
    rng = np.random.default_rng(4)
    ntip = 65536
    parent, leaves = synth.yule_parent(ntip, rng)
    ref = synth.ape_edges(parent, leaves)[:2]
    moves = int(np.ceil(0.05 * (ntip - 3)))
    reps = [synth.ape_edges(synth.spr(parent, leaves, moves, rng), leaves)[:2] for _ in range(2)]
    args = (ref[0], ref[1], [e for e, _ in reps], [n for _, n in reps], ntip)
    monkeypatch.delenv("CASPHY_TBE", raising=False)
    ymin, ysize, xmin, xsize = hip.transfer_trees(*args)
    monkeypatch.setenv("CASPHY_TBE", "bitsets")
    y2, ys2, x2, xs2 = hip.transfer_trees(*args)
    assert np.array_equal(ymin, y2) and np.array_equal(ysize, ys2)
    assert all(np.array_equal(a, b) for a, b in zip(xmin, x2)) and all(np.array_equal(a, b) for a, b in zip(xsize, xs2))
    Y, _ = orc.clade_bitsets(ref[0], ref[1], ntip)
    X, _ = orc.clade_bitsets(reps[0][0], reps[0][1], ntip)
    sel = rng.choice(Y.shape[0], 6, replace=False)
    yo, yso, _, _ = orc.transfer(Y[sel], X, ntip)
    assert np.array_equal(ymin[0][sel], yo) and np.array_equal(ysize[sel], yso)
    sel = rng.choice(X.shape[0], 6, replace=False)
    xo, xso, _, _ = orc.transfer(X[sel], Y, ntip)
    assert np.array_equal(xmin[0][sel], xo) and np.array_equal(xsize[0][sel], xso)


@pytest.mark.parametrize("rpl,waves", [(1, 1), (2, 4), (4, 1), (8, 1), (8, 4)])
def test_transfer_trees_walk_variants(hip, orc, monkeypatch, rpl, waves):
    """Every clades-per-lane / waves-per-workgroup form of the walk kernel
    (lanes past the last clade included) against the oracle."""
    monkeypatch.delenv("CASPHY_TBE", raising=False)
    monkeypatch.setenv("CASPHY_TBE_RPL", str(rpl))
    monkeypatch.setenv("CASPHY_TBE_WAVES", str(waves))
    rng = np.random.default_rng(23)
    ntip = 777
    parent, leaves = synth.yule_parent(ntip, rng)
    ref = synth.ape_edges(parent, leaves)[:2]
    reps = [synth.ape_edges(synth.spr(parent, leaves, 40, rng), leaves)[:2], synth.random_tree(rng, ntip, 200)]
    ymin, ysize, xmin, xsize = hip.transfer_trees(ref[0], ref[1], [e for e, _ in reps], [n for _, n in reps], ntip)
    for b, (y2, ys2, x2, xs2) in enumerate(brute_trees(orc, ntip, ref, reps)):
        assert np.array_equal(ymin[b], y2) and np.array_equal(ysize, ys2), b
        assert np.array_equal(xmin[b], x2) and np.array_equal(xsize[b], xs2), b
