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
