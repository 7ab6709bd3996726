"""The no-allocation tryRadius entry point (casphy_try_radii_bit_noalloc)
against casphy_try_radii_bit, through the C ABI on device buffers the test
allocates itself: two calls in a row (the second on a smaller input, so the
shared scratch array holds the first call's contents), and an input too
large for the scratch array refused with CASPHY_EARG."""
import ctypes as C

import numpy as np
import pytest

from casphy_amd import synth

pytestmark = pytest.mark.gpu


class Dev:
    """Device buffers through the HIP runtime (the caller's own memory)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 4))) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # H2D
        self.bufs.append(p)
        return p

    def zeros(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        assert self.hip.hipMemset(p, 0, C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0  # D2H
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def noalloc_counts(lib, dev, mat, radii, q):
    n, W = mat.shape
    r = np.ascontiguousarray(radii, np.int32)
    qq = np.ascontiguousarray(q, np.uint32)
    dm, dr, dq = dev.put(mat.astype(np.uint32)), dev.put(r), dev.put(qq)
    dc = dev.zeros(r.size * qq.size * 4)
    rc = lib.casphy_try_radii_bit_noalloc(dm, n, W, dr, r.size, dq, qq.size, dc, None)
    assert rc == 0, lib.casphy_last_error()
    return dev.get(dc, (r.size, qq.size), np.uint32)


def test_try_radii_noalloc_matches_plain(hip):
    from casphy_amd import _lib
    lib = _lib.load()
    dev = Dev()
    try:
        mat = synth.binary_matrix(seed=81, n=3000, nsite=1000, p=0.01)  # 32 words: padded rows
        q = synth.query_rows(82, 3000, 500)
        radii = np.arange(1, 41)
        got = noalloc_counts(lib, dev, mat, radii, q)
        want = hip.try_radii(np.ascontiguousarray(mat.T), radii, q)
        assert np.array_equal(got, want)
        # a smaller second call: fewer cells, queries and radii, narrower rows
        small = mat[:1200, :20]
        q2 = synth.query_rows(83, 1200, 77)
        radii2 = np.arange(3, 11)
        got2 = noalloc_counts(lib, dev, small, radii2, q2)
        want2 = hip.try_radii(np.ascontiguousarray(small.T), radii2, q2)
        assert np.array_equal(got2, want2)
        # too large for the scratch array: refused before any work
        big_n, big_W = 200000, 128
        rc = lib.casphy_try_radii_bit_noalloc(C.c_void_p(1), big_n, big_W, C.c_void_p(1), 4, C.c_void_p(1), 1024,
                                              C.c_void_p(1), None)
        assert rc == -1 and b"MiB of scratch" in lib.casphy_last_error()
    finally:
        dev.free()
