"""Product quantization and distance entry points on the device.

    ChunkOffsets::partition        diskann-quantization/src/views.rs:233-302
    populate_chunk_distances       diskann-providers/src/model/pq/fixed_chunk_pq_table.rs:149-200
    BasicTable::compress_into      diskann-quantization/src/product/tables/basic.rs:161-193
    LightPQTrainingParameters      diskann-quantization/src/product/train.rs:22-231
    DistanceProvider               diskann-vector/src/distance/distance_provider.rs:44-46
"""
import ctypes as C

import numpy as np

from . import _native
from .errors import check


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def chunk_offsets(dim, num_chunks):
    lib = _native.load()
    out = np.zeros(num_chunks + 1, dtype=np.int64)
    check(lib.dkm_pq_chunk_offsets(dim, num_chunks, _p(out)), lib)
    return out


def train_pq(data, num_chunks, num_centers=256, lloyd_iters=5, seed=0):
    """Returns (pivots[num_centers, dim], offsets[num_chunks + 1])."""
    lib = _native.load()
    x = np.ascontiguousarray(data, dtype=np.float32)
    n, dim = x.shape
    pivots = np.zeros((num_centers, dim), dtype=np.float32)
    offsets = np.zeros(num_chunks + 1, dtype=np.int64)
    check(lib.dkm_pq_train(_p(x), n, dim, num_chunks, num_centers, lloyd_iters,
                           C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), _p(pivots), _p(offsets)), lib)
    return pivots, offsets


def train_pq_draws(data, num_chunks, draws, num_centers=256, lloyd_iters=5):
    """PQ training with the caller's k-means++ draws (num_chunks x num_centers uniforms in
    [0, 1)); returns (pivots[num_centers, dim], offsets[num_chunks + 1])."""
    lib = _native.load()
    x = np.ascontiguousarray(data, dtype=np.float32)
    n, dim = x.shape
    u = np.ascontiguousarray(draws, dtype=np.float64).reshape(num_chunks, num_centers)
    pivots = np.zeros((num_centers, dim), dtype=np.float32)
    offsets = np.zeros(num_chunks + 1, dtype=np.int64)
    check(lib.dkm_pq_train_draws(_p(x), n, dim, num_chunks, num_centers, lloyd_iters, _p(u),
                                 _p(pivots), _p(offsets)), lib)
    return pivots, offsets


def train_draws(seed, num_chunks, num_centers=256):
    """The draws dkm_pq_train takes from its SplitMix64 stream for `seed` (chunk m, draw c:
    splitmix64(splitmix64(seed ^ splitmix64(0x5851f42d4c957f2d * (m + 1))) + c) >> 11,
    scaled by 2^-53)."""
    mask = (1 << 64) - 1

    def sm(x):
        x = (x + 0x9e3779b97f4a7c15) & mask
        x = ((x ^ (x >> 30)) * 0xbf58476d1ce4e5b9) & mask
        x = ((x ^ (x >> 27)) * 0x94d049bb133111eb) & mask
        return x ^ (x >> 31)

    out = np.zeros((num_chunks, num_centers), np.float64)
    for m in range(num_chunks):
        s = sm((seed & mask) ^ sm((0x5851f42d4c957f2d * (m + 1)) & mask))
        for c in range(num_centers):
            out[m, c] = (sm((s + c) & mask) >> 11) * (1.0 / 9007199254740992.0)
    return out


def pq_encode(pivots, offsets, x):
    lib = _native.load()
    pivots = np.ascontiguousarray(pivots, dtype=np.float32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    M = offsets.shape[0] - 1
    codes = np.zeros((x.shape[0], M), dtype=np.uint8)
    check(lib.dkm_pq_encode(_p(pivots), pivots.shape[0], pivots.shape[1], _p(offsets), M,
                            _p(x), x.shape[0], _p(codes)), lib)
    return codes


def pq_lut(metric, pivots, offsets, queries):
    lib = _native.load()
    pivots = np.ascontiguousarray(pivots, dtype=np.float32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    q = np.ascontiguousarray(queries, dtype=np.float32)
    if q.ndim == 1:
        q = q[None, :]
    M = offsets.shape[0] - 1
    K = pivots.shape[0]
    lut = np.zeros((q.shape[0], M, K), dtype=np.float32)
    check(lib.dkm_pq_lut(int(metric), _p(pivots), K, pivots.shape[1], _p(offsets), M, _p(q),
                         q.shape[0], _p(lut)), lib)
    return lut


def pq_distance(metric, pivots, offsets, codes, ids, queries=None, query_codes=None):
    """Direct PQ distances (pq::distance::DistanceComputer, dynamic.rs:170-205): row i of
    the result holds d(left_i, codes[ids[i, j]]) for the full-precision query rows
    (`queries`) or the code rows `query_codes` (code vs code), exactly one of the two."""
    lib = _native.load()
    pivots = np.ascontiguousarray(pivots, dtype=np.float32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    if ids.ndim == 1:
        ids = ids[None, :]
    nq, m = ids.shape
    M = offsets.shape[0] - 1
    q = None if queries is None else np.ascontiguousarray(queries, dtype=np.float32)
    qc = None if query_codes is None else np.ascontiguousarray(query_codes, dtype=np.uint8)
    out = np.zeros((nq, m), dtype=np.float32)
    check(lib.dkm_pq_distance_batch(int(metric), _p(pivots), pivots.shape[0], pivots.shape[1],
                                    _p(offsets), M, _p(q), _p(qc), nq, _p(codes), codes.shape[0],
                                    _p(ids), m, _p(out)), lib)
    return out


def distance_batch(metric, queries, base, ids):
    """out[i, j] = distance(queries[i], base[ids[i, j]]) in the reference's f32 order."""
    lib = _native.load()
    q = np.ascontiguousarray(queries, dtype=np.float32)
    b = np.ascontiguousarray(base, dtype=np.float32)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    out = np.zeros(ids.shape, dtype=np.float32)
    check(lib.dkm_distance_batch(int(metric), q.shape[1], _p(q), q.shape[0], _p(b), b.shape[0],
                                 _p(ids), ids.shape[1], _p(out)), lib)
    return out


F32 = 0   # k-means assignment in the reference's FP32 arithmetic
BF16 = 1  # k-means assignment as a BF16 MFMA GEMM with FP32 accumulation
FP8 = 2   # candidate pre-scoring as an FP8 (E4M3) MFMA GEMM with FP32 accumulation


class _Dev:
    def __init__(self, lib, arr=None, nbytes=None):
        self.lib = lib
        self.nbytes = arr.nbytes if arr is not None else nbytes
        self.ptr = C.c_void_p()
        check(lib.dkm_device_alloc(max(self.nbytes, 1), C.byref(self.ptr)), lib)
        if arr is not None:
            check(lib.dkm_memcpy_h2d(self.ptr, C.c_void_p(arr.ctypes.data), self.nbytes), lib)

    def get(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        check(self.lib.dkm_memcpy_d2h(C.c_void_p(out.ctypes.data), self.ptr, out.nbytes), self.lib)
        return out

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes != self.nbytes:
            raise ValueError("expected %d bytes, got %d" % (self.nbytes, arr.nbytes))
        check(self.lib.dkm_memcpy_h2d(self.ptr, C.c_void_p(arr.ctypes.data), arr.nbytes), self.lib)

    def __del__(self):
        try:
            self.lib.dkm_device_free(self.ptr)
        except Exception:
            pass


def kmeans(data, init, iters, precision=F32):
    """Lloyd iterations from the given initial centres (lloyds.rs): returns (centres,
    assignment, assign_ms, update_ms). precision F32 reproduces the reference arithmetic;
    BF16 computes the assignment's dot products with bf16 MFMA."""
    lib = _native.load()
    x = np.ascontiguousarray(data, dtype=np.float32)
    c = np.ascontiguousarray(init, dtype=np.float32)
    n, dim = x.shape
    K = c.shape[0]
    dx, dc = _Dev(lib, x), _Dev(lib, c)
    da = _Dev(lib, nbytes=4 * n)
    ta, tu = C.c_double(), C.c_double()
    check(lib.dkm_kmeans_device(dx.ptr, n, dim, K, int(iters), int(precision), dc.ptr, da.ptr,
                                C.byref(ta), C.byref(tu)), lib)
    return dc.get(np.float32, (K, dim)), da.get(np.uint32, n), ta.value, tu.value


def kmeans_assign(data, centres, precision=F32):
    """Nearest centre per row and its score (the FP32 score includes x.x; the BF16 score
    is c.c - 2 x.c)."""
    lib = _native.load()
    x = np.ascontiguousarray(data, dtype=np.float32)
    c = np.ascontiguousarray(centres, dtype=np.float32)
    n, dim = x.shape
    dx, dc = _Dev(lib, x), _Dev(lib, c)
    da, dd = _Dev(lib, nbytes=4 * n), _Dev(lib, nbytes=4 * n)
    check(lib.dkm_kmeans_assign_device(dx.ptr, n, dim, dc.ptr, c.shape[0], int(precision), da.ptr,
                                       dd.ptr), lib)
    return da.get(np.uint32, n), dd.get(np.float32, n)


def closest_centers(data, centers, k=1):
    """compute_closest_centers of the partitioned build: for each row of `data`, the ids of
    its k closest `centers` (ascending squared L2 in the |x|^2 + |c|^2 - 2<x,c> form)."""
    lib = _native.load()
    x = np.ascontiguousarray(data, dtype=np.float32)
    c = np.ascontiguousarray(centers, dtype=np.float32)
    if x.ndim != 2 or c.ndim != 2 or x.shape[1] != c.shape[1]:
        raise ValueError("data and centers must be matrices of the same dimension")
    out = np.zeros((x.shape[0], int(k)), dtype=np.uint32)
    check(lib.dkm_closest_centers(_p(x), x.shape[0], x.shape[1], _p(c), c.shape[0], int(k),
                                  _p(out)), lib)
    return out
