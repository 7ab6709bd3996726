"""nodeKeep on the GPU against the oracle's restatement of cu/nodeKeepBit.cu /
cu/nodeKeepIndel.cu, on trees from the library's own joins."""
import numpy as np
import pytest

from casphy_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nsite", [(3, 32), (150, 256), (700, 1024)])
def test_node_keep_bit(hip, orc, n, nsite):
    mat = synth.binary_matrix(seed=n, n=n, nsite=nsite, p=0.02)
    eo, ns2 = hip.join_edge_bit_reverse(mat)
    rng = np.random.default_rng(n)
    for nout in (0, 1, 17, 300):
        nout = min(nout, n)
        out = mat[rng.choice(n, nout, replace=False)] if nout else np.zeros((0, mat.shape[1]), np.uint32)
        for cutoff in (0, 1, 3, 8):
            got = hip.node_keep(eo, ns2, out, cutoff)
            assert np.array_equal(got, orc.node_keep(eo, ns2, out, cutoff)), (nout, cutoff)


def test_node_keep_indel(hip, orc):
    mat = synth.indel_matrix(seed=8, depth=8, nsite=120)
    n, ncol = mat.shape
    w = np.ones(int(mat.max()) * ncol)
    eo, ns2 = hip.join_edge_indel(mat + 1, w, int(mat.max()), None, 0, 2.0)
    rng = np.random.default_rng(3)
    for nout in (0, 5, 130):
        out = (mat[rng.choice(n, nout, replace=False)] + 1).astype(np.uint32)
        for cutoff in (1, 2, 6):
            got = hip.node_keep(eo, ns2, out, cutoff, indel=True)
            assert np.array_equal(got, orc.node_keep(eo, ns2, out, cutoff, indel=True)), (nout, cutoff)
