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


def test_transfer_trees(hip, orc):
    """Trees in, transfer distances out: equal to the batched transfer over the
    trees' clade bitsets and to the oracle's transfer on the oracle's bitsets."""
    rng = np.random.default_rng(44)
    ntip = 3000
    parent, leaves = synth.yule_parent(ntip, rng)
    ref, rn, _ = synth.ape_edges(parent, leaves)
    reps = []
    for k in (5, 60, 300):
        e, n, _ = synth.ape_edges(synth.spr(parent, leaves, k, rng), leaves)
        reps.append((e, n))
    ymin, ysize, xm, xs = hip.transfer_trees(ref, rn, [e for e, _ in reps], [n for _, n in reps], ntip)
    Y, _ = orc.clade_bitsets(ref, rn, ntip)
    Xs = [orc.clade_bitsets(e, n, ntip)[0] for e, n in reps]
    yb, ysb, xmb, xsb = hip.transfer_dis_batch(Y, Xs, ntip)
    assert np.array_equal(ymin, yb) and np.array_equal(ysize, ysb)
    for b, X in enumerate(Xs):
        assert np.array_equal(xm[b], xmb[b]) and np.array_equal(xs[b], xsb[b])
        y2, ys2, x2, xs2 = orc.transfer(Y, X, ntip)
        assert np.array_equal(ymin[b], y2) and np.array_equal(xm[b], x2) and np.array_equal(xs[b], xs2)
