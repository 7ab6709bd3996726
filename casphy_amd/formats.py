"""CasPhy on-disk formats (read/write), unchanged from the reference.

* ``[u32 dim0, u32 dim1]`` header + raw little-endian 32-bit payload: written by
  write.mat2bin (R/CasPhy_functions.R:1551-1583) as ``dim = rev(dim(mat))`` and
  column-major data, read back by read.bin2mat (:1585-1605).  Used by .tbmat,
  .bmat, .tipSize, .edge2, .nodeState2, .nNeighbor, _xmin.bin, _ymin.bin and by
  fasta2bin's .bin (cu/fasta2bin.cu:335-346).
* headerless payloads: ``.w`` (f64 weights, writeBin(as.numeric(w))) and
  ``.idx`` (concatenated u32 neighbour lists, cu/radiusNeighbor.cu:397).

A file with header (d0, d1) holds d0 records of d1 values; ``read_bin`` returns
it as a (d0, d1) C-order array, i.e. the same bytes.
"""
from __future__ import annotations

import numpy as np


def write_bin(path: str, arr: np.ndarray, dtype=np.uint32) -> None:
    a = np.ascontiguousarray(arr)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    d0, d1 = a.shape
    with open(path, "wb") as f:
        f.write(np.array([d0, d1], dtype="<u4").tobytes())
        f.write(a.astype(np.dtype(dtype).newbyteorder("<"), copy=False).tobytes())


def read_bin(path: str, dtype=np.uint32) -> np.ndarray:
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(8), dtype="<u4")
        if hdr.size != 2:
            raise ValueError(f"{path}: truncated header")
        d0, d1 = int(hdr[0]), int(hdr[1])
        data = np.frombuffer(f.read(), dtype=np.dtype(dtype).newbyteorder("<"))
    if data.size < d0 * d1:
        raise ValueError(f"{path}: payload holds {data.size} values, header says {d0}x{d1}")
    return data[: d0 * d1].reshape(d0, d1).astype(dtype)


def write_weights(path: str, w) -> None:
    np.ascontiguousarray(w, dtype="<f8").tofile(path)


def read_weights(path: str, count: int) -> np.ndarray:
    w = np.fromfile(path, dtype="<f8")
    if w.size < count:
        raise ValueError(f"{path}: {w.size} weights, {count} needed")
    return w[:count].astype(np.float64)


def write_idx(path: str, idx) -> None:
    np.ascontiguousarray(idx, dtype="<u4").tofile(path)


def read_idx(path: str) -> np.ndarray:
    return np.fromfile(path, dtype="<u4").astype(np.uint32)
