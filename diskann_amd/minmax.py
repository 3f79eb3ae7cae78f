"""8-bit MinMax quantization on the device (diskann-quantization/src/minmax).

    MinMaxQuantizer::compress         minmax/quantizer.rs:45-170
    FullQuery compression             minmax/quantizer.rs:178-210
    MinMax{IP, L2Squared, Cosine, CosineNormalized}   minmax/vectors.rs:206-470

A code row is u32 dim, f32 b, f32 n, f32 a, f32 |reconstruction|^2, then the codes.
"""
import ctypes as C

import numpy as np

from . import _native
from .errors import check
from .spherical import TRANSFORMS


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class MinMaxQuantizer:
    def __init__(self, dim, grid_scale=1.0, transform="null", signs0=None, signs1=None,
                 padded_dim=0, subsample=None):
        self.lib = _native.load()
        self.dim = int(dim)
        s0 = None if signs0 is None else np.ascontiguousarray(signs0, dtype=np.uint8)
        s1 = None if signs1 is None else np.ascontiguousarray(signs1, dtype=np.uint8)
        sub = None if subsample is None else np.ascontiguousarray(subsample, dtype=np.uint32)
        self._h = C.c_void_p()
        check(self.lib.dkm_minmax_create(self.dim, float(grid_scale), TRANSFORMS[transform],
                                         _p(s0), _p(s1), 0 if s1 is None else s1.shape[0],
                                         int(padded_dim), _p(sub),
                                         0 if sub is None else sub.shape[0], C.byref(self._h)),
              self.lib)
        od, cb = C.c_int64(), C.c_int64()
        check(self.lib.dkm_minmax_dims(self._h, C.byref(od), C.byref(cb)), self.lib)
        self.out_dim, self.code_bytes = od.value, cb.value

    def compress(self, x, with_loss=False):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        out = np.zeros((x.shape[0], self.code_bytes), np.uint8)
        loss = np.zeros(x.shape[0], np.float32)
        check(self.lib.dkm_minmax_compress(self._h, _p(x), x.shape[0], _p(out), _p(loss)),
              self.lib)
        return (out, loss) if with_loss else out

    def distance(self, metric, left, right, ids):
        left = np.ascontiguousarray(np.atleast_2d(left), dtype=np.uint8)
        right = np.ascontiguousarray(np.atleast_2d(right), dtype=np.uint8)
        ids = np.ascontiguousarray(np.atleast_2d(ids), dtype=np.uint32)
        out = np.zeros(ids.shape, np.float32)
        check(self.lib.dkm_minmax_distance_batch(self._h, int(metric), _p(left), left.shape[0],
                                                 _p(right), right.shape[0], _p(ids), ids.shape[1],
                                                 _p(out)), self.lib)
        return out

    def query_distance(self, metric, queries, codes, ids):
        q = np.ascontiguousarray(np.atleast_2d(queries), dtype=np.float32)
        codes = np.ascontiguousarray(np.atleast_2d(codes), dtype=np.uint8)
        ids = np.ascontiguousarray(np.atleast_2d(ids), dtype=np.uint32)
        out = np.zeros(ids.shape, np.float32)
        check(self.lib.dkm_minmax_query_distance_batch(self._h, int(metric), _p(q), q.shape[0],
                                                       _p(codes), codes.shape[0], _p(ids),
                                                       ids.shape[1], _p(out)), self.lib)
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.dkm_minmax_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
