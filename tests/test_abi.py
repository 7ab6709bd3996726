"""CPU checks of the C ABI: the library loads and exports every declared symbol,
and host-side argument validation fails loudly without touching a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            src = open(os.path.join(inc, f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(casphy_\w+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    from casphy_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared surface
    assert {n for n, _, _ in _lib.SIGNATURES} == names


def test_version_and_error_channel():
    from casphy_amd import _lib
    L = _lib.load()
    assert L.casphy_version() >= 10000
    assert isinstance(L.casphy_last_error(), bytes)


def test_bad_arguments_rejected_before_gpu_work():
    """NULL buffers and n < 2 are argument errors (no device needed)."""
    from casphy_amd import _lib
    L = _lib.load()
    edge = np.zeros(4, np.uint32)
    rc = L.casphy_join_bit(None, 3, 1, None, 1, edge.ctypes.data_as(C.c_void_p), None, None)
    assert rc == -1
    assert b"bad arguments" in L.casphy_last_error()
    mat = np.zeros((1, 1), np.uint32)
    rc = L.casphy_join_bit(mat.ctypes.data_as(C.c_void_p), 1, 1, None, 1, edge.ctypes.data_as(C.c_void_p), None, None)
    assert rc == -1 and b"two cells" in L.casphy_last_error()
    rc = L.casphy_join_bit(mat.ctypes.data_as(C.c_void_p), 2, 1, None, 7, edge.ctypes.data_as(C.c_void_p), None, None)
    assert rc == -1 and b"ancestralType" in L.casphy_last_error()


def test_formats_roundtrip(tmp_path):
    from casphy_amd import formats
    a = np.arange(12, dtype=np.uint32).reshape(3, 4)
    p = str(tmp_path / "x.tbmat")
    formats.write_bin(p, a)
    raw = np.fromfile(p, dtype="<u4")
    assert raw[0] == 3 and raw[1] == 4 and np.array_equal(raw[2:], a.ravel())
    assert np.array_equal(formats.read_bin(p), a)
    with open(p, "r+b") as f:
        f.truncate(20)
    with pytest.raises(ValueError):
        formats.read_bin(p)


def test_tip_merge_lists_checked_before_gpu_work():
    """tipMerge refuses malformed lists and over-long lists before any device work."""
    from casphy_amd import _lib
    L = _lib.load()
    vp = C.c_void_p
    member = np.zeros(3, np.int32)

    def call(nnb, idx):
        nnb = np.asarray(nnb, np.uint32)
        idx = np.asarray(idx, np.uint32)
        return L.casphy_tip_merge(nnb.ctypes.data_as(vp), idx.ctypes.data_as(vp), nnb.size, C.c_double(0.5),
                                  member.ctypes.data_as(vp))
    assert call([2, 1, 1], [1, 0, 1, 2]) == -1 and b"not strictly ascending" in L.casphy_last_error()
    assert call([1, 1, 1], [0, 1, 3]) == -1 and b"out of range" in L.casphy_last_error()
    assert call([1 << 20 | 1, 1, 1], [0]) == -1 and b"1,048,576" in L.casphy_last_error()


def test_node_keep_edges_checked_before_gpu_work():
    from casphy_amd import _lib
    L = _lib.load()
    vp = C.c_void_p
    st = np.zeros((2, 1), np.uint32)
    kept = np.zeros(2, np.uint32)
    nk = C.c_uint32(0)

    def call(edges):
        e = np.asarray(edges, np.uint32)
        return L.casphy_node_keep_bit(e.ctypes.data_as(vp), e.shape[0], st.ctypes.data_as(vp), 2, 1, None, 0, 1,
                                      kept.ctypes.data_as(vp), C.byref(nk))
    assert call([[3, 4], [4, 0], [4, 1], [3, 9]]) == -1 and b"beyond the tree" in L.casphy_last_error()
    assert call([[3, 4], [3, 4], [4, 1], [3, 2]]) == -1 and b"child of two edges" in L.casphy_last_error()


def test_clade_bitsets_edge_order_checked():
    """Edge lists where a node's edge comes after its children's are refused
    (nodeDesBig would write partial rows for them)."""
    from casphy_amd import api
    good = np.array([[4, 5], [5, 1], [5, 2], [4, 3]], np.int32)      # 3 tips, root 4
    bad = good[[1, 2, 0, 3]]
    with pytest.raises(api.CasphyError) as e:
        api.clade_bitsets(bad, 2, 3)
    assert e.value.code == -1 and "leaves a node before the edge into it" in str(e.value)
    with pytest.raises(api.CasphyError):
        api.clade_bitsets(np.array([[4, 5], [5, 1], [5, 2], [5, 3], [4, 5]], np.int32), 2, 3)


def test_root_vote_arguments_checked_before_gpu_work():
    from casphy_amd import api
    st = np.zeros((4, 2), np.uint32)
    mat = np.zeros((3, 2), np.uint32)
    good = np.array([[1, 2], [1, 3], [3, 4]], np.uint32)
    for fn in (api.vote_root, api.find_root):
        with pytest.raises(api.CasphyError) as e:
            fn(st, good, mat, [0, 3])
        assert "names row 3 of 3" in str(e.value)
        with pytest.raises(api.CasphyError) as e:
            fn(st, np.array([[1, 2], [1, 5]], np.uint32), mat, [0])
        assert "outside 1..4" in str(e.value)
        with pytest.raises(api.CasphyError) as e:
            fn(st, np.array([[1, 2], [3, 2]], np.uint32), mat, [0])
        assert "child of two edges" in str(e.value)
        with pytest.raises(api.CasphyError) as e:
            fn(np.zeros((0, 2), np.uint32), np.zeros((0, 2), np.uint32), mat, [0])
        assert "no node" in str(e.value)


def test_fasta_encode_arguments_checked():
    from casphy_amd import _lib
    L = _lib.load()
    seqs = np.frombuffer(b"ACGT", np.uint8)
    rc = L.casphy_fasta_encode(seqs.ctypes.data_as(C.c_void_p), 1, 4, None, None, None, None)
    assert rc == -1 and b"bad arguments" in L.casphy_last_error()
