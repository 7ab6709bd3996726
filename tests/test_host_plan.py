"""The library's host-only planning code (casphy_amd/csrc/plan.cpp and
newick.cpp) built with g++ under the address and undefined-behaviour
sanitizers and driven by tests/host_check.cpp: clade-row plans, the
transfer-distance walk programs (replayed on the CPU with the device kernel's
op semantics) and nodeKeep's skip rule, each against the oracle; the Newick
reader against the oracle's reader.  No GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from casphy_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TREES = "/root/reference/examples/test.bstrees"


@pytest.fixture(scope="session")
def host_check(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("hostbin") / "host_check"
    src = [os.path.join(ROOT, "tests", "host_check.cpp"), os.path.join(ROOT, "casphy_amd", "csrc", "plan.cpp"),
           os.path.join(ROOT, "casphy_amd", "csrc", "newick.cpp")]
    subprocess.check_call([gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=undefined", "-o", str(out)] + src)
    return str(out)


def run(exe, cmd, d):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, cmd, str(d)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return r


def walk_case(exe, tmp_path, orc, ntip, ref, reps):
    (ye, yn) = ref
    d = tmp_path / "walk"
    d.mkdir(exist_ok=True)
    with open(d / "meta.txt", "w") as f:
        f.write(f"{ntip} {ye.shape[0]} {yn} {len(reps)}\n")
        for e, n in reps:
            f.write(f"{e.shape[0]} {n}\n")
    np.ascontiguousarray(ye[:, 0], np.int32).tofile(d / "y_anc.bin")
    np.ascontiguousarray(ye[:, 1], np.int32).tofile(d / "y_des.bin")
    xa = np.concatenate([e[:, 0] for e, _ in reps]).astype(np.int32) if reps else np.zeros(0, np.int32)
    xd = np.concatenate([e[:, 1] for e, _ in reps]).astype(np.int32) if reps else np.zeros(0, np.int32)
    xa.tofile(d / "x_anc.bin")
    xd.tofile(d / "x_des.bin")
    run(exe, "walk", d)
    ymin = np.fromfile(d / "ymin.bin", np.uint32)
    ysize = np.fromfile(d / "ysize.bin", np.uint32)
    xmin = np.fromfile(d / "xmin.bin", np.uint32)
    xsize = np.fromfile(d / "xsize.bin", np.uint32)
    depth = int(open(d / "depth.txt").read())
    Y, _ = orc.clade_bitsets(ye, yn, ntip)
    ny = Y.shape[0]
    assert ymin.size == ny * len(reps)
    xo = 0
    for b, (e, n) in enumerate(reps):
        X, _ = orc.clade_bitsets(e, n, ntip)
        y2, ys2, x2, xs2 = orc.transfer(Y, X, ntip)
        assert np.array_equal(ymin[b * ny:(b + 1) * ny], y2), b
        assert np.array_equal(ysize, ys2)
        assert np.array_equal(xmin[xo:xo + X.shape[0]], x2), b
        assert np.array_equal(xsize[xo:xo + X.shape[0]], xs2), b
        xo += X.shape[0]
    assert xo == xmin.size
    return depth


@pytest.mark.parametrize("ntip", [4, 33, 300])
def test_walk_yule_spr(host_check, tmp_path, orc, ntip):
    """Transfer minima of a Yule tree against SPR-perturbed copies (binary
    trees) equal the oracle's brute force over clade bitsets; the walk stack
    stays within log2(ntip) + 1."""
    rng = np.random.default_rng(ntip)
    parent, leaves = synth.yule_parent(ntip, rng)
    ref = synth.ape_edges(parent, leaves)[:2]
    reps = [synth.ape_edges(synth.spr(parent, leaves, max(1, ntip // 10), rng), leaves)[:2] for _ in range(3)]
    reps.append(ref)  # the identical tree: every minimum 0
    depth = walk_case(host_check, tmp_path, orc, ntip, ref, reps)
    assert depth <= int(np.floor(np.log2(ntip))) + 1


def test_walk_multifurcating_and_single_child(host_check, tmp_path, orc):
    rng = np.random.default_rng(5)
    for ntip, nint in ((50, 10), (200, 120), (120, 3)):
        ref = synth.random_tree(rng, ntip, nint)
        reps = [synth.random_tree(rng, ntip, int(rng.integers(1, ntip))) for _ in range(3)]
        depth = walk_case(host_check, tmp_path, orc, ntip, ref, reps)
        assert depth <= int(np.floor(np.log2(ntip))) + 1


def test_walk_caterpillar_and_star(host_check, tmp_path, orc):
    """The deepest tree still walks with two stack values; a star has no clade
    rows, so its replicate minima stay at 32 * ceil(ntip / 32)."""
    ntip = 700
    cat = synth.caterpillar(ntip)
    rng = np.random.default_rng(8)
    parent, leaves = synth.yule_parent(ntip, rng)
    yule = synth.ape_edges(parent, leaves)[:2]
    star = synth.ape_edges(np.array([-1] + [0] * ntip), np.arange(1, ntip + 1))[:2]
    depth = walk_case(host_check, tmp_path, orc, ntip, cat, [yule, cat, star])
    assert depth <= int(np.floor(np.log2(ntip))) + 1
    walk_case(host_check, tmp_path, orc, ntip, yule, [cat])


def test_clade_plan(host_check, tmp_path, orc):
    """Rows rebuilt from the plan (tips order[lo .. lo + cnt)) and the size
    vector equal nodeDesBig's (the oracle)."""
    rng = np.random.default_rng(3)
    for ntip, nint in ((40, 39), (300, 80), (64, 1)):
        edge, nnode = synth.random_tree(rng, ntip, nint)
        d = tmp_path / f"clades{ntip}"
        d.mkdir()
        with open(d / "meta.txt", "w") as f:
            f.write(f"{ntip} {edge.shape[0]} {nnode}\n")
        np.ascontiguousarray(edge[:, 0], np.int32).tofile(d / "anc.bin")
        np.ascontiguousarray(edge[:, 1], np.int32).tofile(d / "des.bin")
        run(host_check, "clades", d)
        order = np.fromfile(d / "order.bin", np.uint32)
        lo = np.fromfile(d / "lo.bin", np.uint32)
        cnt = np.fromfile(d / "cnt.bin", np.uint32)
        size = np.fromfile(d / "size.bin", np.int32)
        rows, sizes = orc.clade_bitsets(edge, nnode, ntip)
        assert sorted(order.tolist()) == list(range(ntip))
        assert lo.size == rows.shape[0] and np.array_equal(size, sizes)
        for r in range(lo.size):
            tips = order[lo[r]:lo[r] + cnt[r]]
            bits = np.zeros(rows.shape[1] * 32, bool)
            bits[tips] = True
            assert np.array_equal(synth.pack_bits(bits[None, :])[0], rows[r]), r


def test_node_keep_walk(host_check, tmp_path, orc):
    """nodeKeep's edge walk with the retain flags of the bit form (every
    outgroup row lacks at least `cutoff` of the node's mutations) equals the
    oracle's nodeKeepBit on joined trees."""
    rng = np.random.default_rng(4)
    for n, nout, cutoff in ((60, 5, 1), (300, 40, 2), (120, 0, 1)):
        mat = synth.binary_matrix(seed=int(n), n=n, nsite=256, p=0.02)
        edge, ns, _ = orc.join_bit(mat)
        eo = np.zeros_like(edge)
        io = np.zeros(n - 1, np.uint32)
        orc.lib().orc_cladewise(np.ascontiguousarray(edge.reshape(-1)), n, eo.reshape(-1), io)
        state = ns[io]  # rows of .nodeState2: internal nodes in relabel order
        out = mat[rng.choice(n, nout, replace=False)] if nout else np.zeros((0, mat.shape[1]), np.uint32)
        bits = synth.unpack_bits(state)
        if nout:
            miss = (bits[:, None, :] & ~synth.unpack_bits(out)[None, :, :]).sum(2)
            retain = (miss >= cutoff).all(1).astype(np.uint32)
        else:
            retain = np.ones(n - 1, np.uint32)
        d = tmp_path / f"nk{n}"
        d.mkdir()
        with open(d / "meta.txt", "w") as f:
            f.write(f"{eo.shape[0]} {n}\n")
        np.ascontiguousarray(eo, np.uint32).tofile(d / "edge.bin")
        retain.tofile(d / "retain.bin")
        run(host_check, "nodekeep", d)
        kept = np.fromfile(d / "kept.bin", np.uint32)
        assert np.array_equal(kept, orc.node_keep(eo, state, out, cutoff)), n


def test_newick_reader(host_check, tmp_path, orc):
    """The library's reader (under the sanitizers) on written trees with
    lengths, labels and multifurcations, and on the reference's bootstrap
    trees when present, equals the oracle's reader."""
    rng = np.random.default_rng(9)
    texts = []
    for ntip in (2, 17, 400):
        parent, leaves = synth.yule_parent(ntip, rng)
        texts.append(synth.newick(parent, leaves, rng=rng))
    texts.append(synth.newick(np.array([-1] + [0] * 30), np.arange(1, 31)))
    texts.append("((A:1,B:2)x:3,(C,'D''q')[c]:1,E);")
    cases = [("\n".join(texts)).encode()]
    if os.path.exists(REF_TREES):
        with open(REF_TREES, "rb") as f:
            cases.append(f.read(4 << 20).rsplit(b";", 1)[0] + b";")  # whole trees of the first 4 MiB
    for k, text in enumerate(cases):
        d = tmp_path / f"nwk{k}"
        d.mkdir()
        (d / "text.nwk").write_bytes(text)
        run(host_check, "newick", d)
        meta = [tuple(map(int, line.split())) for line in open(d / "trees.txt")]
        anc = np.fromfile(d / "anc.bin", np.int32)
        des = np.fromfile(d / "des.bin", np.int32)
        want = orc.read_newick(text)
        assert len(meta) == len(want)
        o = 0
        for (nt, nn, ne), (edge, nnode, ntip, _) in zip(meta, want):
            assert (nt, nn, ne) == (ntip, nnode, edge.shape[0])
            assert np.array_equal(anc[o:o + ne], edge[:, 0]) and np.array_equal(des[o:o + ne], edge[:, 1])
            o += ne
