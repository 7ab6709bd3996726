"""Fixtures built from the reference's own example inputs (tests/golden/,
made by tests/golden/make_golden.py from data/mat1k.rda, data/mat_binary.rda
and examples/run1_1024.fasta).  CPU tests re-derive the fixtures' expected
values with the oracle and with independent numpy closed forms; GPU tests
hold the HIP library to the same bytes."""
import os

import numpy as np
import pytest

from casphy_amd import synth
from tests import brute

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(G, name), allow_pickle=False))


def dna_scores(tb):
    b = tb.view(np.uint8).reshape(tb.shape[0], -1) & 0x0F
    return ((b[:, None, :] & b[None, :, :]) != 0).sum(axis=2)


# ------------------------------------------------------------------ CPU pins
def test_binary_fixture_oracle_and_closed_form(orc):
    f = load("binary_mat.npz")
    S = brute.score_matrix_bits(f["mat"])
    for r in f["radii"]:
        c, i = brute.radius_lists_from_matrix(S, int(r))
        assert np.array_equal(f[f"counts_r{r}"], c)
        if f"idx_r{r}" in f:
            assert np.array_equal(f[f"idx_r{r}"], i)
    e, ns, _ = orc.join_bit(f["mat"], ancestral_type=1, lazy=True)
    assert np.array_equal(e, f["edge"]) and np.array_equal(ns, f["node_state"])
    eo, order = brute.cladewise_R(f["edge"], f["mat"].shape[0])
    assert np.array_equal(eo, f["edge2"]) and np.array_equal(order, f["order"])


def test_indel_fixture_oracle(orc):
    f = load("indel_mat1k.npz")
    mat = f["mat"]
    assert mat.min() >= -1 and np.array_equal(synth.regularize(mat), mat)
    w = np.ones((mat.max(), mat.shape[1]), np.float32)
    e, ns, br = orc.join_indel(mat, w, lazy=True)
    assert np.array_equal(e, f["edge"]) and np.array_equal(ns, f["node_state"])
    assert np.array_equal(br, f["branch"])
    # tips keep their states; every internal state follows the indel rule of its children
    assert np.array_equal(ns[:mat.shape[0]], mat)


def test_dna_fixture_closed_form():
    f = load("dna_run1.npz")
    S = dna_scores(f["mat"])
    assert np.array_equal(S, S.T)
    for r in f["radii"]:
        c, i = brute.radius_lists_from_matrix(S, int(r))
        assert np.array_equal(f[f"counts_r{r}"], c) and np.array_equal(f[f"idx_r{r}"], i)
        assert np.array_equal(f[f"try_r{r}"], c[f["q"]])


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_binary_fixture_gpu(hip):
    f = load("binary_mat.npz")
    bmat = np.ascontiguousarray(f["mat"].T)
    for r in f["radii"]:
        c, i = hip.radius_neighbor(bmat, int(r))
        assert np.array_equal(c, f[f"counts_r{r}"])
        if f"idx_r{r}" in f:
            assert np.array_equal(i, f[f"idx_r{r}"])
    e, ns, _ = hip.join_bit(f["mat"])
    assert np.array_equal(e, f["edge"]) and np.array_equal(ns, f["node_state"])
    eo, order = hip.join_edge_bit_reverse(f["mat"])
    assert np.array_equal(eo, f["edge2"]) and np.array_equal(order, f["node_state"][f["order"]])


@pytest.mark.gpu
def test_indel_fixture_gpu(hip):
    f = load("indel_mat1k.npz")
    e, ns, br = hip.join_indel(f["mat"])
    assert np.array_equal(e, f["edge"]) and np.array_equal(ns, f["node_state"])
    assert np.array_equal(br, f["branch"])
    out = hip.greedyJoinEdge2(f["mat"], np.ones((f["mat"].max(), 30)), np.ones(1000, np.uint32))
    assert np.array_equal(out[:-1, 0:2], f["edge"]) and np.array_equal(out[:-1, 2], f["branch"])


@pytest.mark.gpu
def test_dna_fixture_gpu(hip):
    f = load("dna_run1.npz")
    for r in f["radii"]:
        c, i = hip.radius_neighbor_dna(f["mat"], int(r))
        assert np.array_equal(c, f[f"counts_r{r}"]) and np.array_equal(i, f[f"idx_r{r}"])
    got = hip.try_radii_dna(f["mat"], f["radii"], f["q"])
    for k, r in enumerate(f["radii"]):
        assert np.array_equal(got[k], f[f"try_r{r}"])
