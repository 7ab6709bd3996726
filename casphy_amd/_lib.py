"""ctypes binding of libcasphy_hip.so (the HIP product).

The library must have been built in-tree (``make -C casphy_amd`` or
``__graft_entry__.build()``).  There is no CPU fallback: when the library is
missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcasphy_hip.so")

u32, i32, u64, f32 = C.c_uint32, C.c_int32, C.c_uint64, C.c_float
vp = C.c_void_p

# (name, restype, argtypes) of every symbol declared in include/casphy_hip.h
SIGNATURES = [
    ("casphy_last_error", C.c_char_p, []),
    ("casphy_version", C.c_int, []),
    ("casphy_free", None, [vp]),
    ("casphy_set_device", C.c_int, [C.c_int]),
    ("casphy_synchronize", C.c_int, []),
    ("casphy_mfma_tuning", C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("casphy_score_bit_matrix", C.c_int, [vp, u32, u32, vp, vp]),
    ("casphy_score_indel_matrix", C.c_int, [vp, u32, u32, vp, u32, vp]),
    ("casphy_radius_bit", C.c_int, [vp, u32, u32, vp, i32, vp, C.POINTER(C.POINTER(u32)), C.POINTER(u64)]),
    ("casphy_radius_dna", C.c_int, [vp, u32, u32, i32, vp, C.POINTER(C.POINTER(u32)), C.POINTER(u64)]),
    ("casphy_try_radius_bit", C.c_int, [vp, u32, u32, vp, i32, vp, u32, vp]),
    ("casphy_try_radius_dna", C.c_int, [vp, u32, u32, i32, vp, u32, vp]),
    ("casphy_try_radii_bit", C.c_int, [vp, u32, u32, vp, vp, u32, vp, u32, vp]),
    ("casphy_try_radii_dna", C.c_int, [vp, u32, u32, vp, u32, vp, u32, vp]),
    ("casphy_try_radii_bit_noalloc", C.c_int, [vp, u32, u32, vp, u32, vp, u32, vp, vp]),
    ("casphy_join_bit", C.c_int, [vp, u32, u32, vp, C.c_int, vp, vp, vp]),
    ("casphy_join_indel", C.c_int, [vp, u32, u32, vp, u32, vp, u32, f32, vp, vp, vp]),
    ("casphy_cladewise", C.c_int, [vp, u32, vp, vp]),
    ("casphy_join_bit_batch", C.c_int, [vp, vp, u32, u32, vp, C.c_int, vp]),
    ("casphy_join_indel_batch", C.c_int, [vp, vp, u32, u32, vp, u32, vp, u32, f32, vp]),
    ("casphy_join_indel_bootstrap", C.c_int, [vp, u32, u32, vp, u32, vp]),
    ("casphy_transfer", C.c_int, [vp, u32, vp, u32, u32, u32, vp, vp, vp, vp]),
    ("casphy_transfer_batch", C.c_int, [vp, u32, vp, vp, u32, u32, u32, vp, vp, vp, vp]),
    ("casphy_tip_merge", C.c_int, [vp, vp, u32, C.c_double, vp]),
    ("casphy_node_keep_bit", C.c_int, [vp, u32, vp, u32, u32, vp, u32, u32, vp, C.POINTER(u32)]),
    ("casphy_clade_bitsets", C.c_int, [vp, vp, u32, u32, u32, vp, C.POINTER(u32), vp]),
    ("casphy_transfer_trees", C.c_int, [vp, vp, u32, u32, vp, vp, vp, vp, u32, u32, vp, vp, vp, vp]),
    ("casphy_node_keep_indel", C.c_int, [vp, u32, vp, u32, u32, vp, u32, u32, vp, C.POINTER(u32)]),
    ("casphy_vote_root", C.c_int, [vp, u32, u32, vp, u32, vp, u64, vp, u32, vp, vp]),
    ("casphy_find_root", C.c_int, [vp, u32, u32, vp, u32, vp, u64, vp, u32, vp, vp]),
    ("casphy_fasta_encode", C.c_int, [vp, u32, u64, vp, vp, vp, C.POINTER(u64)]),
    ("casphy_newick_read", C.c_int, [vp, u64, vp, u64, C.POINTER(u64), C.POINTER(u32), C.POINTER(u32),
                                     C.POINTER(vp), C.POINTER(u32), C.POINTER(vp), C.POINTER(u64)]),
    ("casphy_tbe_support", C.c_int, [vp, u32, u32, vp, u32, vp]),
    ("casphy_cells_create", C.c_int, [C.c_int, vp, u32, u32, C.c_int, vp, C.POINTER(vp)]),
    ("casphy_cells_destroy", None, [vp]),
    ("casphy_cells_radius", C.c_int, [vp, i32, C.POINTER(u64)]),
    ("casphy_cells_fetch", C.c_int, [vp, vp, vp]),
    ("casphy_cells_tip_merge", C.c_int, [vp, C.c_double, vp]),
    ("casphy_cells_split_groups", C.c_int, [vp, vp, vp, u32, vp, u32, vp, vp, C.c_double, vp, vp]),
    ("casphy_cells_group_fetch", C.c_int, [vp, u32, vp, vp]),
    ("casphy_cells_group_tip_merge", C.c_int, [vp, u32, C.c_double, vp]),
    ("casphy_cells_try_radii", C.c_int, [vp, vp, u32, vp, u32, vp]),
    ("casphy_cells_autosplit", C.c_int, [vp, vp, u32, vp, u32, C.c_double, C.POINTER(i32), C.POINTER(u64), vp]),
    ("casphy_cells_last_ms", C.c_int, [vp, C.POINTER(f32), C.POINTER(f32)]),
    ("casphy_server_create", C.c_int, [C.c_int, u64, u32, C.POINTER(vp)]),
    ("casphy_server_destroy", None, [vp]),
    ("casphy_server_submit_join", C.c_int, [vp, C.c_int, vp, u32, u32, C.c_int, C.POINTER(u64)]),
    ("casphy_server_submit_transfer", C.c_int, [vp, vp, u32, vp, u32, u32, u32, C.POINTER(u64)]),
    ("casphy_server_put_matrix", C.c_int, [vp, C.c_int, vp, u32, u32, C.POINTER(u64)]),
    ("casphy_server_drop_matrix", C.c_int, [vp, u64]),
    ("casphy_server_submit_join_rows", C.c_int, [vp, u64, vp, u32, C.c_int, C.POINTER(u64)]),
    ("casphy_server_submit_transfer_trees", C.c_int, [vp, vp, vp, u32, u32, vp, vp, vp, vp, u32, u32, C.POINTER(u64)]),
    ("casphy_server_wait", C.c_int, [vp, u64, vp, u64, C.POINTER(C.c_double)]),
    ("casphy_server_stats", C.c_int, [vp, C.POINTER(u64), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(u64)]),
]

_lib = None


class CasphyError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"casphy error {code}: {msg}")
        self.code = code


def load():
    """Loads the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C casphy_amd` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().casphy_last_error()
        raise CasphyError(rc, msg.decode() if msg else "")
