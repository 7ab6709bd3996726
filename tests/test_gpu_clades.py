"""Clade bitsets (nodeDesBig) on the GPU against the oracle's restatement."""
import numpy as np
import pytest

from casphy_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ntip", [2, 3, 1000, 65536])
def test_clade_bitsets_yule(hip, orc, ntip):
    parent, leaves = synth.yule_parent(ntip, np.random.default_rng(ntip))
    edge, nnode, nt = synth.ape_edges(parent, leaves)
    rows, sizes = hip.clade_bitsets(edge, nnode, nt)
    r2, s2 = orc.clade_bitsets(edge, nnode, nt)
    assert np.array_equal(rows, r2) and np.array_equal(sizes, s2)


def test_clade_bitsets_caterpillar_and_star(hip, orc):
    # caterpillar: depth ntip - 1
    ntip = 3000
    parent = np.full(2 * ntip - 1, -1, np.int64)
    spine = [0]
    leaves = []
    nxt = 1
    for k in range(ntip - 1):
        a, b = nxt, nxt + 1
        nxt += 2
        parent[a] = parent[b] = spine[-1]
        leaves.append(a)
        spine.append(b)
    leaves.append(spine[-1])
    edge, nnode, nt = synth.ape_edges(parent, np.array(leaves))
    rows, sizes = hip.clade_bitsets(edge, nnode, nt)
    r2, s2 = orc.clade_bitsets(edge, nnode, nt)
    assert np.array_equal(rows, r2) and np.array_equal(sizes, s2)
    # two wide internal nodes under the root, rows of 18,750 words (no LDS row)
    ntip = 600000
    half = ntip // 2
    parent = np.concatenate([[-1, 0, 0], np.full(half, 1), np.full(ntip - half, 2)]).astype(np.int64)
    leaves = np.arange(3, 3 + ntip)
    edge, nnode, nt = synth.ape_edges(parent, leaves)
    rows, sizes = hip.clade_bitsets(edge, nnode, nt)
    r2, s2 = orc.clade_bitsets(edge, nnode, nt)
    assert rows.shape == (2, 18750) and np.array_equal(rows, r2) and np.array_equal(sizes, s2)
