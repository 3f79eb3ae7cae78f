"""The reference's on-disk formats through the C ABI (host-side file I/O).

    write_bin / read_bin            diskann-utils/src/io.rs:60-100
    write_pq_pivots / read_pq_pivots diskann-providers/src/storage/pq_storage.rs:93-149,278-349
    graph files                     DiskANNIndex.save_graph / save_graph_disk / load_graph
                                    (diskann-providers/src/storage/bin.rs:230-380)
"""
import ctypes as C

import numpy as np

from . import _native
from .errors import check


def _path(p):
    return str(p).encode()


def write_bin(path, array):
    """Writes a 2-D array as a .bin matrix (u32 rows, u32 cols, row-major data)."""
    a = np.ascontiguousarray(array)
    if a.ndim != 2:
        raise ValueError("write_bin needs a 2-D array")
    lib = _native.load()
    check(lib.dkm_bin_write(_path(path), C.c_void_p(a.ctypes.data), a.dtype.itemsize,
                            a.shape[0], a.shape[1]), lib)


def read_bin_shape(path):
    lib = _native.load()
    r, c = C.c_int64(), C.c_int64()
    check(lib.dkm_bin_read_shape(_path(path), C.byref(r), C.byref(c)), lib)
    return r.value, c.value


def read_bin(path, dtype=np.float32):
    """Reads a .bin matrix of `dtype`."""
    lib = _native.load()
    r, c = read_bin_shape(path)
    out = np.empty((r, c), dtype=dtype)
    check(lib.dkm_bin_read(_path(path), C.c_void_p(out.ctypes.data), out.dtype.itemsize, r, c),
          lib)
    return out


def write_pq_pivots(path, pivots, offsets, centroid=None):
    """Writes a PQ pivots file (pivots K x D, chunk offsets M + 1, optional centroid)."""
    p = np.ascontiguousarray(pivots, dtype=np.float32)
    o = np.ascontiguousarray(offsets, dtype=np.int64)
    cen = None if centroid is None else np.ascontiguousarray(centroid, dtype=np.float32)
    if cen is not None and cen.shape != (p.shape[1],):
        raise ValueError("centroid must have the pivots' dimension")
    lib = _native.load()
    check(lib.dkm_pq_pivots_write(_path(path), C.c_void_p(p.ctypes.data), p.shape[0], p.shape[1],
                                  None if cen is None else C.c_void_p(cen.ctypes.data),
                                  C.c_void_p(o.ctypes.data), o.shape[0] - 1), lib)


def read_pq_pivots(path):
    """Reads a PQ pivots file: (pivots K x D with a non-zero centroid added, offsets)."""
    lib = _native.load()
    K, D, M = C.c_int64(), C.c_int64(), C.c_int64()
    check(lib.dkm_pq_pivots_read_shape(_path(path), C.byref(K), C.byref(D), C.byref(M)), lib)
    piv = np.empty((K.value, D.value), np.float32)
    off = np.empty(M.value + 1, np.int64)
    check(lib.dkm_pq_pivots_read(_path(path), C.c_void_p(piv.ctypes.data), K.value, D.value,
                                 C.c_void_p(off.ctypes.data), M.value), lib)
    return piv, off


def write_pq_codes(path, codes):
    write_bin(path, np.ascontiguousarray(codes, dtype=np.uint8))


def read_pq_codes(path):
    return read_bin(path, np.uint8)
