"""ctypes binding for oracle/liboracle.so (test infrastructure only)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

COSINE, INNER_PRODUCT, L2, COSINE_NORMALIZED = 0, 1, 2, 3
MODE_FULL, MODE_HYBRID, MODE_PQ_ONLY = 0, 1, 2

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _setup(_lib)
    return _lib


class Index(C.Structure):
    _fields_ = [
        ("metric", C.c_int), ("dim", C.c_int64), ("vstride", C.c_int64),
        ("vectors", C.c_void_p), ("max_degree", C.c_int64), ("adj", C.c_void_p),
        ("deg", C.c_void_p), ("nstart", C.c_int64), ("starts", C.c_void_p),
        ("M", C.c_int64), ("K", C.c_int64), ("codes", C.c_void_p),
        ("pivots", C.c_void_p), ("offsets", C.c_void_p), ("sq", C.c_void_p),
        ("sph", C.c_void_p), ("sph_codes", C.c_void_p), ("sph_layout", C.c_int),
        ("sph_norescale", C.c_int),
    ]


class SQ(C.Structure):
    _fields_ = [
        ("dim", C.c_int64), ("scale", C.c_float), ("shift", C.c_void_p),
        ("shift_square_norm", C.c_float), ("mean_norm", C.c_float), ("codes", C.c_void_p),
        ("comp", C.c_void_p),
    ]


class Stats(C.Structure):
    _fields_ = [("cmps", C.c_uint32), ("hops", C.c_uint32), ("count", C.c_uint32)]


class BuildCfg(C.Structure):
    _fields_ = [
        ("metric", C.c_int), ("R", C.c_int64), ("R_max", C.c_int64),
        ("L_build", C.c_int64), ("alpha", C.c_float), ("max_occlusion", C.c_int64),
        ("saturate", C.c_int), ("max_backedges", C.c_int64),
        ("max_fp", C.c_int64), ("codes", C.c_void_p), ("pivots", C.c_void_p),
        ("pq_offsets", C.c_void_p), ("pq_M", C.c_int64), ("sq", C.c_void_p),
        ("sph", C.c_void_p), ("sph_codes", C.c_void_p),
    ]


class Graph(C.Structure):
    _fields_ = [("rows", C.c_int64), ("max_degree", C.c_int64), ("adj", C.c_void_p),
                ("deg", C.c_void_p)]


def _setup(L):
    P, I64, F = C.c_void_p, C.c_int64, C.c_float
    L.dko_distance.argtypes = [C.c_int, P, P, I64]
    L.dko_distance.restype = F
    L.dko_square_norm.argtypes = [P, I64]
    L.dko_square_norm.restype = F
    L.dko_chunk_offsets.argtypes = [I64, I64, P]
    L.dko_pq_lut.argtypes = [C.c_int, P, I64, I64, P, I64, P, P]
    L.dko_pq_lut_transposed.argtypes = [C.c_int, P, I64, I64, P, I64, P, P]
    L.dko_closest_centers.argtypes = [P, I64, I64, P, I64, I64, P]
    L.dko_disk_search.argtypes = [C.POINTER(Index), P, I64, I64, I64, C.c_uint64, P, P,
                                  C.POINTER(Stats)]
    L.dko_pq_adc.argtypes = [P, P, I64, I64]
    L.dko_pq_adc.restype = F
    L.dko_sq_train.argtypes = [P, I64, I64, I64, C.c_double, P, P, P]
    L.dko_sq_shift_square_norm.argtypes = [P, I64]
    L.dko_sq_shift_square_norm.restype = F
    L.dko_sq_compress.argtypes = [F, P, I64, P, P, P]
    L.dko_sq_compress.restype = C.c_int
    L.dko_sq_rescale.argtypes = [P, I64, F]
    L.dko_sq_distance.argtypes = [C.c_int, F, F, P, F, P, F, I64]
    L.dko_sq_distance.restype = F
    S = C.POINTER(SphStruct)
    L.dko_sph_train_stats.argtypes = [C.c_int, P, I64, I64, P, P]
    L.dko_sph_data_bytes.argtypes = [S]
    L.dko_sph_data_bytes.restype = I64
    L.dko_sph_hadamard.argtypes = [P, I64]
    L.dko_sph_transform.argtypes = [S, P, P]
    L.dko_sph_rescale.argtypes = [S, P]
    L.dko_sph_compress_data.argtypes = [S, P, P]
    L.dko_sph_compress_data.restype = C.c_int
    L.dko_sph_compress_query4.argtypes = [S, P, P, P]
    L.dko_sph_compress_query4.restype = C.c_int
    L.dko_sph_compress_full.argtypes = [S, P, P, P]
    L.dko_sph_compress_full.restype = C.c_int
    L.dko_sph_dist_data.argtypes = [S, P, P]
    L.dko_sph_dist_data.restype = F
    L.dko_sph_dist_query4.argtypes = [S, P, P, P]
    L.dko_sph_dist_query4.restype = F
    L.dko_sph_dist_full.argtypes = [S, P, P, P]
    L.dko_sph_dist_full.restype = F
    L.dko_minmax_bytes.argtypes = [S]
    L.dko_minmax_bytes.restype = I64
    L.dko_minmax_compress.argtypes = [S, F, P, P, P]
    L.dko_minmax_compress.restype = C.c_int
    L.dko_minmax_compress_full.argtypes = [S, P, P, P]
    L.dko_minmax_compress_full.restype = C.c_int
    L.dko_minmax_dist_data.argtypes = [C.c_int, P, P]
    L.dko_minmax_dist_data.restype = F
    L.dko_minmax_dist_full.argtypes = [C.c_int, P, P, P]
    L.dko_minmax_dist_full.restype = F
    L.dko_f32_to_f16.argtypes = [F]
    L.dko_f32_to_f16.restype = C.c_uint16
    L.dko_f16_to_f32.argtypes = [C.c_uint16]
    L.dko_f16_to_f32.restype = F
    L.dko_pq_encode.argtypes = [P, I64, I64, P, I64, P, I64, P]
    L.dko_pq_direct.argtypes = [C.c_int, P, I64, P, I64, P, P]
    L.dko_pq_direct.restype = C.c_float
    L.dko_pq_qq.argtypes = [C.c_int, P, I64, P, I64, P, P]
    L.dko_pq_qq.restype = C.c_float
    L.dko_pq_encode.restype = C.c_int
    L.dko_knn_validate.argtypes = [I64, I64, I64]
    L.dko_search.argtypes = [C.POINTER(Index), C.c_int, P, I64, I64, I64, P, P,
                             C.POINTER(Stats)]
    L.dko_search_full.argtypes = [C.POINTER(Index), C.c_int, P, I64, I64, P, P, P, P, P, P,
                                  I64, C.POINTER(Stats)]
    L.dko_search_batch.argtypes = [C.POINTER(Index), C.c_int, P, I64, I64, I64, I64, I64,
                                   C.c_int, P, P, P, P, P]
    L.dko_search_batch.restype = C.c_double
    L.dko_search_batch_mixed.argtypes = [C.POINTER(Index), C.c_int, P, I64, I64, P, P, I64,
                                         I64, C.c_int, P, P, P, P, P]
    L.dko_search_batch_mixed.restype = C.c_double
    L.dko_occlude_list.argtypes = [C.POINTER(BuildCfg), P, I64, I64, P, P, I64, C.c_uint32, P]
    L.dko_occlude_list.restype = I64
    L.dko_prune_list.argtypes = [C.POINTER(BuildCfg), P, I64, I64, C.c_uint32, P, I64, C.c_int, P]
    L.dko_prune_list.restype = I64
    L.dko_update_occlude_factor.argtypes = [C.c_int, F, F, F, F]
    L.dko_update_occlude_factor.restype = F
    L.dko_build.argtypes = [C.POINTER(BuildCfg), P, I64, I64, C.POINTER(Graph), P, I64, P,
                            I64, I64, I64]
    L.dko_medoid.argtypes = [P, I64, I64, I64, P]
    L.dko_medoid.restype = I64
    L.dko_lloyd_assign.argtypes = [P, I64, I64, P, I64, P]
    L.dko_lloyd_assign.restype = F
    L.dko_lloyd_update.argtypes = [P, I64, I64, P, I64, P]
    L.dko_kmeanspp.argtypes = [P, I64, I64, I64, I64, P, P, P]
    L.dko_kmeanspp.restype = C.c_int
    L.dko_pq_train.argtypes = [P, I64, I64, P, I64, I64, I64, P, P]
    L.dko_pq_train.restype = C.c_int
    L.dko_bruteforce_f64.argtypes = [C.c_int, P, I64, I64, I64, P, I64, I64, C.c_int, P, P]


def _p(a):
    return None if a is None else a.ctypes.data


def distance(metric, q, x):
    q = np.ascontiguousarray(q, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    return lib().dko_distance(metric, _p(q), _p(x), q.shape[0])


def square_norm(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return lib().dko_square_norm(_p(x), x.shape[0])


def chunk_offsets(dim, m):
    out = np.zeros(m + 1, dtype=np.int64)
    lib().dko_chunk_offsets(dim, m, _p(out))
    return out


def pq_lut(metric, pivots, offsets, query):
    pivots = np.ascontiguousarray(pivots, dtype=np.float32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    query = np.ascontiguousarray(query, dtype=np.float32)
    K, dim = pivots.shape
    M = offsets.shape[0] - 1
    lut = np.zeros(M * K, dtype=np.float32)
    lib().dko_pq_lut(metric, _p(pivots), K, dim, _p(offsets), M, _p(query), _p(lut))
    return lut.reshape(M, K)


def closest_centers(x, c, k=1):
    x = np.ascontiguousarray(x, dtype=np.float32)
    c = np.ascontiguousarray(c, dtype=np.float32)
    out = np.zeros((x.shape[0], k), dtype=np.uint32)
    lib().dko_closest_centers(_p(x), x.shape[0], x.shape[1], _p(c), c.shape[0], k, _p(out))
    return out


def pq_lut_transposed(metric, pivots, offsets, query):
    pivots = np.ascontiguousarray(pivots, dtype=np.float32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    query = np.ascontiguousarray(query, dtype=np.float32)
    M = offsets.shape[0] - 1
    lut = np.zeros(M * pivots.shape[0], dtype=np.float32)
    lib().dko_pq_lut_transposed(metric, _p(pivots), pivots.shape[0], pivots.shape[1],
                                _p(offsets), M, _p(query), _p(lut))
    return lut


def pq_adc(code, lut):
    code = np.ascontiguousarray(code, dtype=np.uint8)
    lut = np.ascontiguousarray(lut, dtype=np.float32)
    return lib().dko_pq_adc(_p(code), _p(lut), lut.shape[0], lut.shape[1])


def pq_direct(metric, pivots, offsets, query, code):
    """Full-precision query vs the pivots a code selects (fixed_chunk_pq_table.rs:32-56)."""
    L = lib()
    pv = np.ascontiguousarray(pivots, dtype=np.float32)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    q = np.ascontiguousarray(query, dtype=np.float32)
    c = np.ascontiguousarray(code, dtype=np.uint8)
    return float(L.dko_pq_direct(int(metric), _p(pv), pv.shape[1], _p(off), len(off) - 1, _p(q),
                                 _p(c)))


def pq_qq(metric, pivots, offsets, a, b):
    """Code vs code through the pivot table (fixed_chunk_pq_table.rs:290-365)."""
    L = lib()
    pv = np.ascontiguousarray(pivots, dtype=np.float32)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    ca = np.ascontiguousarray(a, dtype=np.uint8)
    cb = np.ascontiguousarray(b, dtype=np.uint8)
    return float(L.dko_pq_qq(int(metric), _p(pv), pv.shape[1], _p(off), len(off) - 1, _p(ca),
                             _p(cb)))


def pq_encode(pivots, offsets, x):
    pivots = np.ascontiguousarray(pivots, dtype=np.float32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    K, dim = pivots.shape
    M = offsets.shape[0] - 1
    codes = np.zeros((x.shape[0], M), dtype=np.uint8)
    rc = lib().dko_pq_encode(_p(pivots), K, dim, _p(offsets), M, _p(x), x.shape[0], _p(codes))
    if rc:
        raise ValueError("PQ encode: infinite or NaN distance at row %d" % (rc - 1))
    return codes


def sq_train(x, stdevs=2.0):
    """ScalarQuantizationParameters::train: (scale, shift[dim], mean_norm)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, dim = x.shape
    shift = np.zeros(dim, np.float32)
    scale, mn = C.c_float(), C.c_float()
    lib().dko_sq_train(_p(x), n, dim, dim, float(stdevs), C.byref(scale), _p(shift),
                       C.byref(mn))
    return float(np.float32(scale.value)), shift, float(np.float32(mn.value))


def sq_shift_square_norm(shift):
    shift = np.ascontiguousarray(shift, dtype=np.float32)
    return lib().dko_sq_shift_square_norm(_p(shift), shift.shape[0])


def sq_compress(scale, shift, x):
    """8-bit codes [n, dim] and compensations [n] of the rows of x; raises on NaN."""
    shift = np.ascontiguousarray(shift, dtype=np.float32)
    x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
    codes = np.zeros(x.shape, np.uint8)
    comp = np.zeros(x.shape[0], np.float32)
    for r in range(x.shape[0]):
        c = C.c_float()
        if lib().dko_sq_compress(scale, _p(shift), x.shape[1], _p(x[r]), _p(codes[r]),
                                 C.byref(c)):
            raise ValueError("input contains NaN (row %d)" % r)
        comp[r] = c.value
    return codes, comp


def sq_rescale(x, to_norm):
    x = np.array(x, dtype=np.float32)
    lib().dko_sq_rescale(_p(x), x.shape[0], to_norm)
    return x


def sq_distance(metric, scale, shift_square_norm, a, comp_a, b, comp_b):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    return lib().dko_sq_distance(metric, scale, shift_square_norm, _p(a), comp_a, _p(b),
                                 comp_b, a.shape[0])


class SphStruct(C.Structure):
    _fields_ = [
        ("metric", C.c_int), ("dim", C.c_int64), ("out_dim", C.c_int64), ("shift", C.c_void_p),
        ("mean_norm", C.c_float), ("pre_scale", C.c_float), ("kind", C.c_int),
        ("signs0", C.c_void_p), ("signs1", C.c_void_p), ("padded_dim", C.c_int64),
        ("subsample", C.c_void_p), ("nsub", C.c_int64),
    ]


SPH_NULL, SPH_PADDING_HADAMARD, SPH_DOUBLE_HADAMARD = 0, 1, 2


def _signs(b):
    return np.ascontiguousarray(np.where(np.asarray(b, bool), 0x80000000, 0), dtype=np.uint32)


class Spherical:
    """1-bit spherical quantizer on the oracle. kind: SPH_*; signs0 / signs1 booleans
    (True = flip); subsample: optional index array."""

    def __init__(self, metric, shift, mean_norm, pre_scale=1.0, kind=SPH_DOUBLE_HADAMARD,
                 signs0=None, signs1=None, padded_dim=0, subsample=None):
        self.shift = np.ascontiguousarray(shift, dtype=np.float32)
        dim = self.shift.shape[0]
        self.signs0 = _signs(signs0 if signs0 is not None else np.zeros(dim, bool))
        self.signs1 = _signs(signs1) if signs1 is not None else None
        self.subsample = (None if subsample is None else
                          np.ascontiguousarray(subsample, dtype=np.uint32))
        if self.subsample is not None:
            out = self.subsample.shape[0]
        elif kind == SPH_PADDING_HADAMARD:
            out = padded_dim
        elif kind == SPH_DOUBLE_HADAMARD:
            out = self.signs1.shape[0]
        else:
            out = dim
        self.out_dim = out
        self.metric = metric
        self.s = SphStruct(metric, dim, out, _p(self.shift), mean_norm, pre_scale, kind,
                           _p(self.signs0), _p(self.signs1), padded_dim, _p(self.subsample),
                           0 if self.subsample is None else self.subsample.shape[0])
        self.data_bytes = lib().dko_sph_data_bytes(C.byref(self.s))

    def transform(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros(self.out_dim, np.float32)
        lib().dko_sph_transform(C.byref(self.s), _p(x), _p(out))
        return out

    def rescale(self, x):
        x = np.array(x, dtype=np.float32)
        lib().dko_sph_rescale(C.byref(self.s), _p(x))
        return x

    def compress(self, x):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        out = np.zeros((x.shape[0], self.data_bytes), np.uint8)
        for r in range(x.shape[0]):
            rc = lib().dko_sph_compress_data(C.byref(self.s), _p(x[r]), _p(out[r]))
            if rc:
                raise ValueError("spherical compression error %d (row %d)" % (rc, r))
        return out

    def compress_query4(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        codes = np.zeros(self.out_dim, np.uint8)
        meta = np.zeros(4, np.float32)
        rc = lib().dko_sph_compress_query4(C.byref(self.s), _p(x), _p(codes), _p(meta))
        if rc:
            raise ValueError("spherical compression error %d" % rc)
        return codes, meta

    def compress_full(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        v = np.zeros(self.out_dim, np.float32)
        meta = np.zeros(3, np.float32)
        rc = lib().dko_sph_compress_full(C.byref(self.s), _p(x), _p(v), _p(meta))
        if rc:
            raise ValueError("spherical compression error %d" % rc)
        return v, meta

    def dist_data(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint8)
        b = np.ascontiguousarray(b, dtype=np.uint8)
        return lib().dko_sph_dist_data(C.byref(self.s), _p(a), _p(b))

    def dist_query4(self, codes, meta, b):
        b = np.ascontiguousarray(b, dtype=np.uint8)
        return lib().dko_sph_dist_query4(C.byref(self.s), _p(codes), _p(meta), _p(b))

    def dist_full(self, v, meta, b):
        b = np.ascontiguousarray(b, dtype=np.uint8)
        return lib().dko_sph_dist_full(C.byref(self.s), _p(v), _p(meta), _p(b))


def sph_train_stats(metric, x):
    """(centroid, mean_norm) of SphericalQuantizer::train."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    c = np.zeros(x.shape[1], np.float32)
    mn = C.c_float()
    lib().dko_sph_train_stats(metric, _p(x), x.shape[0], x.shape[1], _p(c), C.byref(mn))
    return c, float(np.float32(mn.value))


class MinMax:
    """8-bit MinMax quantizer on the oracle; the transform comes from a Spherical's."""

    def __init__(self, transform, grid_scale=1.0):
        self.t = transform
        self.grid_scale = float(grid_scale)
        self.bytes = lib().dko_minmax_bytes(C.byref(transform.s))

    def compress(self, x):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        out = np.zeros((x.shape[0], self.bytes), np.uint8)
        loss = np.zeros(x.shape[0], np.float32)
        for r in range(x.shape[0]):
            lo = C.c_float()
            if lib().dko_minmax_compress(C.byref(self.t.s), self.grid_scale, _p(x[r]), _p(out[r]),
                                         C.byref(lo)):
                raise ValueError("input contains NaN (row %d)" % r)
            loss[r] = lo.value
        return out, loss

    def compress_full(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        v = np.zeros(self.t.out_dim, np.float32)
        m = np.zeros(2, np.float32)
        if lib().dko_minmax_compress_full(C.byref(self.t.s), _p(x), _p(v), _p(m)):
            raise ValueError("input contains NaN")
        return v, m

    @staticmethod
    def dist_data(metric, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint8)
        b = np.ascontiguousarray(b, dtype=np.uint8)
        return lib().dko_minmax_dist_data(metric, _p(a), _p(b))

    @staticmethod
    def dist_full(metric, v, m, b):
        b = np.ascontiguousarray(b, dtype=np.uint8)
        return lib().dko_minmax_dist_full(metric, _p(v), _p(m), _p(b))


def hadamard(x):
    x = np.array(x, dtype=np.float32)
    lib().dko_sph_hadamard(_p(x), x.shape[0])
    return x


def f32_to_f16_bits(v):
    return lib().dko_f32_to_f16(float(v))


class OracleIndex:
    """Holds numpy arrays alive behind an Index struct. sq = (scale, shift, mean_norm,
    codes, comp) selects scalar-quantized traversal for the quantized modes."""

    def __init__(self, metric, vectors, adj, deg, starts, codes=None, pivots=None,
                 offsets=None, sq=None, sph=None):
        self.vectors = np.ascontiguousarray(vectors, dtype=np.float32)
        self.adj = np.ascontiguousarray(adj, dtype=np.uint32)
        self.deg = np.ascontiguousarray(deg, dtype=np.uint32)
        self.starts = np.ascontiguousarray(starts, dtype=np.uint32)
        s = Index()
        s.metric = metric
        s.dim = self.vectors.shape[1]
        s.vstride = self.vectors.shape[1]
        s.vectors = _p(self.vectors)
        s.max_degree = self.adj.shape[1]
        s.adj = _p(self.adj)
        s.deg = _p(self.deg)
        s.nstart = self.starts.shape[0]
        s.starts = _p(self.starts)
        if codes is not None:
            self.codes = np.ascontiguousarray(codes, dtype=np.uint8)
            self.pivots = np.ascontiguousarray(pivots, dtype=np.float32)
            self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
            s.M = self.codes.shape[1]
            s.K = self.pivots.shape[0]
            s.codes = _p(self.codes)
            s.pivots = _p(self.pivots)
            s.offsets = _p(self.offsets)
        if sq is not None:
            scale, shift, mean_norm, sqc, comp = sq
            self.sq_shift = np.ascontiguousarray(shift, dtype=np.float32)
            self.sq_codes = np.ascontiguousarray(sqc, dtype=np.uint8)
            self.sq_comp = np.ascontiguousarray(comp, dtype=np.float32)
            self.sq = SQ(self.vectors.shape[1], scale, _p(self.sq_shift),
                         sq_shift_square_norm(self.sq_shift), mean_norm,
                         _p(self.sq_codes), _p(self.sq_comp))
            s.sq = C.cast(C.pointer(self.sq), C.c_void_p)
        if sph is not None:
            # Spherical, canonical rows, layout 0 / 1 [, no inner-product query rescale]
            quant, sph_codes, layout = sph[:3]
            s.sph_norescale = int(len(sph) > 3 and bool(sph[3]))
            self.sph_q = quant
            self.sph_codes = np.ascontiguousarray(sph_codes, dtype=np.uint8)
            s.sph = C.cast(C.pointer(quant.s), C.c_void_p)
            s.sph_codes = _p(self.sph_codes)
            s.sph_layout = layout
        self.s = s

    def search(self, query, k, L, bw=1, mode=MODE_FULL):
        query = np.ascontiguousarray(query, dtype=np.float32)
        ids = np.zeros(k, dtype=np.uint32)
        ds = np.zeros(k, dtype=np.float32)
        st = Stats()
        rc = lib().dko_search(C.byref(self.s), mode, _p(query), k, L, bw, _p(ids), _p(ds),
                              C.byref(st))
        if rc:
            raise ValueError("knn validation error %d" % rc)
        n = st.count
        return ids[:n], ds[:n], st.cmps, st.hops

    def search_disk(self, query, k, L, bw=1, io_limit=None):
        """Disk-index search from the single start row (the medoid)."""
        query = np.ascontiguousarray(query, dtype=np.float32)
        ids = np.zeros(k, dtype=np.uint32)
        ds = np.zeros(k, dtype=np.float32)
        st = Stats()
        lim = (1 << 64) - 1 if io_limit is None else int(io_limit)
        rc = lib().dko_disk_search(C.byref(self.s), _p(query), k, L, bw, lim, _p(ids), _p(ds),
                                   C.byref(st))
        if rc:
            raise ValueError("knn validation error %d" % rc)
        n = st.count
        return ids[:n], ds[:n], st.cmps, st.hops

    def search_full(self, query, L, bw=1, mode=MODE_FULL, rec_cap=1 << 16):
        query = np.ascontiguousarray(query, dtype=np.float32)
        qi = np.zeros(L + self.starts.shape[0], dtype=np.uint32)
        qd = np.zeros(L + self.starts.shape[0], dtype=np.float32)
        ri = np.zeros(rec_cap, dtype=np.uint32)
        rd = np.zeros(rec_cap, dtype=np.float32)
        ql = C.c_uint32()
        rl = C.c_uint32()
        st = Stats()
        lib().dko_search_full(C.byref(self.s), mode, _p(query), L, bw, _p(qi), _p(qd),
                              C.byref(ql), _p(ri), _p(rd), C.byref(rl), rec_cap, C.byref(st))
        n = min(rl.value, rec_cap)
        return qi[:ql.value], qd[:ql.value], ri[:n], rd[:n], st.cmps, st.hops

    def search_batch(self, queries, k, L, bw=1, mode=MODE_FULL, nthreads=1, latency=False):
        queries = np.ascontiguousarray(queries, dtype=np.float32)
        nq = queries.shape[0]
        ids = np.zeros((nq, k), dtype=np.uint32)
        ds = np.zeros((nq, k), dtype=np.float32)
        cm = np.zeros(nq, dtype=np.uint32)
        hp = np.zeros(nq, dtype=np.uint32)
        lat = np.zeros(nq, dtype=np.float64) if latency else None
        wall = lib().dko_search_batch(C.byref(self.s), mode, _p(queries), nq,
                                      queries.shape[1], k, L, bw, nthreads, _p(ids), _p(ds),
                                      _p(cm), _p(hp), _p(lat))
        if wall < 0:
            raise ValueError("knn validation error")
        return ids, ds, cm, hp, wall, lat

    def search_batch_mixed(self, queries, ks, Ls, bw=1, mode=MODE_FULL, nthreads=1):
        """One k and one L per query; rows of ids / ds have max(ks) entries (unused
        entries are 0xFFFFFFFF / inf)."""
        queries = np.ascontiguousarray(queries, dtype=np.float32)
        nq = queries.shape[0]
        ks = np.ascontiguousarray(ks, dtype=np.uint32)
        Ls = np.ascontiguousarray(Ls, dtype=np.uint32)
        kmax = int(ks.max()) if nq else 1
        ids = np.zeros((nq, kmax), dtype=np.uint32)
        ds = np.zeros((nq, kmax), dtype=np.float32)
        cm = np.zeros(nq, dtype=np.uint32)
        hp = np.zeros(nq, dtype=np.uint32)
        wall = lib().dko_search_batch_mixed(C.byref(self.s), mode, _p(queries), nq,
                                            queries.shape[1], _p(ks), _p(Ls), kmax, bw,
                                            nthreads, _p(ids), _p(ds), _p(cm), _p(hp), None)
        if wall < 0:
            raise ValueError("knn validation error")
        return ids, ds, cm, hp, wall


def _sq_struct(sq, dim, keep):
    """dko_sq for sq = (scale, shift, mean_norm, codes (rows x dim), compensations)."""
    scale, shift, mean_norm, sqc, comp = sq
    sh = np.ascontiguousarray(shift, dtype=np.float32)
    cd = np.ascontiguousarray(sqc, dtype=np.uint8)
    cp = np.ascontiguousarray(comp, dtype=np.float32)
    st = SQ(dim, scale, _p(sh), sq_shift_square_norm(sh), mean_norm, _p(cd), _p(cp))
    keep += [sh, cd, cp, st]
    return C.cast(C.pointer(st), C.c_void_p)


def _sph_args(sph, keep):
    """(dko_sph pointer, canonical code rows pointer) for sph = (Spherical, codes)."""
    if sph is None:
        return None, None
    quant, codes = sph
    cd = np.ascontiguousarray(codes, dtype=np.uint8)
    keep += [quant, cd]
    return C.cast(C.pointer(quant.s), C.c_void_p), _p(cd)


def build_graph(metric, vectors, R, R_max, L_build, alpha, starts, ids, batch=1,
                intra_batch=-1, start_adj=None, max_occlusion=750, saturate=False,
                max_backedges=None, max_fp=None, codes=None, pivots=None, pq_offsets=None,
                sq=None, sph=None):
    """Vamana build on the oracle. Returns (adj, deg) for all rows of `vectors`.
    max_fp (with codes / pivots / pq_offsets): Hybrid prune with that many full-precision
    ids per prune; None: full precision throughout. sq = (scale, shift, mean_norm, codes,
    compensations) of every row: the SQ Quantized build (SQ insert search, prune distances
    between stored codes). sph = (Spherical, canonical code rows of every row): the
    spherical Quantized::build() strategy."""
    vectors = np.ascontiguousarray(vectors, dtype=np.float32)
    rows = vectors.shape[0]
    adj = np.zeros((rows, R_max), dtype=np.uint32)
    deg = np.zeros(rows, dtype=np.uint32)
    if start_adj is not None:
        for s, lst in start_adj.items():
            adj[s, :len(lst)] = lst
            deg[s] = len(lst)
    g = Graph(rows, R_max, _p(adj), _p(deg))
    keep = []
    if max_fp is None:
        hyb = (-1, None, None, None, 0)
    else:
        cd = np.ascontiguousarray(codes, dtype=np.uint8)
        pv = np.ascontiguousarray(pivots, dtype=np.float32)
        off = np.ascontiguousarray(pq_offsets, dtype=np.int64)
        keep = [cd, pv, off]
        hyb = (int(max_fp), _p(cd), _p(pv), _p(off), off.shape[0] - 1)
    cfg = BuildCfg(metric, R, R_max, L_build, alpha, max_occlusion, int(saturate),
                   R_max if max_backedges is None else max_backedges, *hyb,
                   _sq_struct(sq, vectors.shape[1], keep) if sq is not None else None,
                   *_sph_args(sph, keep))
    starts = np.ascontiguousarray(starts, dtype=np.uint32)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    lib().dko_build(C.byref(cfg), _p(vectors), vectors.shape[1], vectors.shape[1],
                    C.byref(g), _p(starts), starts.shape[0], _p(ids), ids.shape[0], batch,
                    intra_batch)
    del keep
    return adj, deg


def occlude_list(metric, vectors, pool_ids, pool_d, self_id, R, alpha, saturate=False):
    vectors = np.ascontiguousarray(vectors, dtype=np.float32)
    pool_ids = np.ascontiguousarray(pool_ids, dtype=np.uint32)
    pool_d = np.ascontiguousarray(pool_d, dtype=np.float32)
    cfg = BuildCfg(metric, R, R, 0, alpha, 750, int(saturate), R, -1)
    out = np.zeros(max(R, 1) + pool_ids.shape[0], dtype=np.uint32)
    n = lib().dko_occlude_list(C.byref(cfg), _p(vectors), vectors.shape[1], vectors.shape[1],
                               _p(pool_ids), _p(pool_d), pool_ids.shape[0], self_id, _p(out))
    return out[:n]


def prune_list(metric, vectors, location, lst, R, alpha, max_occlusion=750, saturate=False,
               force_saturate=False, max_fp=None, codes=None, pivots=None, pq_offsets=None,
               sq=None, sph=None):
    """robust_prune_list of `lst` around row `location`; returns the pruned ids (sq as in
    build_graph: distances between stored SQ codes)."""
    vectors = np.ascontiguousarray(vectors, dtype=np.float32)
    lst = np.ascontiguousarray(lst, dtype=np.uint32)
    keep = []
    if max_fp is None:
        hyb = (-1, None, None, None, 0)
    else:
        cd = np.ascontiguousarray(codes, dtype=np.uint8)
        pv = np.ascontiguousarray(pivots, dtype=np.float32)
        off = np.ascontiguousarray(pq_offsets, dtype=np.int64)
        keep = [cd, pv, off]
        hyb = (int(max_fp), _p(cd), _p(pv), _p(off), off.shape[0] - 1)
    cfg = BuildCfg(metric, R, R, 0, alpha, max_occlusion, int(saturate), R, *hyb,
                   _sq_struct(sq, vectors.shape[1], keep) if sq is not None else None,
                   *_sph_args(sph, keep))
    out = np.zeros(max(R, 1), dtype=np.uint32)
    n = lib().dko_prune_list(C.byref(cfg), _p(vectors), vectors.shape[1], vectors.shape[1],
                             int(location), _p(lst), lst.shape[0], int(force_saturate), _p(out))
    del keep
    return out[:n]


def update_occlude_factor(metric, d_ik, d_jk, cur, alpha):
    return lib().dko_update_occlude_factor(metric, d_ik, d_jk, cur, alpha)


def medoid(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    mean = np.zeros(x.shape[1], dtype=np.float32)
    i = lib().dko_medoid(_p(x), x.shape[0], x.shape[1], x.shape[1], _p(mean))
    return i, mean


def lloyd_assign(x, c):
    x = np.ascontiguousarray(x, dtype=np.float32)
    c = np.ascontiguousarray(c, dtype=np.float32)
    nearest = np.zeros(x.shape[0], dtype=np.uint32)
    r = lib().dko_lloyd_assign(_p(x), x.shape[0], x.shape[1], _p(c), c.shape[0], _p(nearest))
    return nearest, r


def lloyd_update(x, nearest, K):
    x = np.ascontiguousarray(x, dtype=np.float32)
    nearest = np.ascontiguousarray(nearest, dtype=np.uint32)
    c = np.zeros((K, x.shape[1]), dtype=np.float32)
    lib().dko_lloyd_update(_p(x), x.shape[0], x.shape[1], _p(nearest), K, _p(c))
    return c


KPP_OK, KPP_DATASET_TOO_SMALL, KPP_INSUFFICIENT_DIVERSITY, KPP_SAW_INFINITY = 0, 1, 2, 3


def kmeanspp(x, K, u):
    """k-means++ seeding with the uniform draws u (K doubles in [0, 1)); returns
    (centers K x ds, selected, reason)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = np.ascontiguousarray(u, dtype=np.float64)
    assert u.shape[0] >= max(K, 1)
    c = np.zeros((K, x.shape[1]), dtype=np.float32)
    sel = C.c_int64(0)
    r = lib().dko_kmeanspp(_p(x), x.shape[0], x.shape[1], x.shape[1], K, _p(u), _p(c),
                           C.byref(sel))
    return c, sel.value, r


def pq_train(x, offsets, K, iters, u):
    """PQ pivots (K x dim) from the train driver with draws u (M x K doubles); raises on
    an unrecoverable k-means++ failure."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    M = off.shape[0] - 1
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(M, K)
    piv = np.zeros((K, x.shape[1]), dtype=np.float32)
    r = lib().dko_pq_train(_p(x), x.shape[0], x.shape[1], _p(off), M, K, iters, _p(u), _p(piv))
    if r:
        raise RuntimeError("k-means++ failed on chunk %d" % (r - 1))
    return piv


def bruteforce_f64(metric, base, queries, k, nthreads=8):
    base = np.ascontiguousarray(base, dtype=np.float32)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    nq = queries.shape[0]
    ids = np.zeros((nq, k), dtype=np.uint32)
    ds = np.zeros((nq, k), dtype=np.float64)
    lib().dko_bruteforce_f64(metric, _p(base), base.shape[0], base.shape[1], base.shape[1],
                             _p(queries), nq, k, nthreads, _p(ids), _p(ds))
    return ids, ds
