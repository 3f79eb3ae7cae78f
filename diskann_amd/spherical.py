"""1-bit spherical quantization on the device (diskann-quantization/src/spherical).

    SphericalQuantizer (train / generate)   spherical/quantizer.rs:36-320
    Impl<1>::compress, distance_computer    spherical/iface.rs:1040-1966
    fused_query_computer (query layouts)    spherical/iface.rs:1898-1950
    serialized form                         schemas/spherical_v001.fbs, transforms_v001.fbs

A compressed vector is the reference's canonical row: ceil(out_dim / 8) bytes of bits,
then f16 inner-product correction, f16 metric term and u16 bit count.
"""
import ctypes as C
import struct

import numpy as np

from . import _native
from .errors import check

METRICS = {"squared_l2": 2, "inner_product": 1, "cosine": 0}
TRANSFORMS = {"null": 0, "padding_hadamard": 1, "double_hadamard": 2}
LAYOUTS = {"same_as_data": 0, "four_bit_transposed": 1, "full_precision": 2}
_FB_METRIC = {0: "squared_l2", 1: "inner_product", 2: "cosine"}
_FB_TRANSFORM = {1: "random_rotation", 2: "padding_hadamard", 3: "double_hadamard", 4: "null"}


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class _Table:
    """Minimal flatbuffer table reader (vtable lookup, scalars, sub-tables, vectors)."""

    def __init__(self, buf, pos):
        self.buf, self.pos = buf, pos
        self.vt = pos - struct.unpack_from("<i", buf, pos)[0]
        self.vt_len = struct.unpack_from("<H", buf, self.vt)[0]

    def _off(self, i):
        o = 4 + 2 * i
        return 0 if o >= self.vt_len else struct.unpack_from("<H", self.buf, self.vt + o)[0]

    def scalar(self, i, fmt, default=0):
        o = self._off(i)
        return default if o == 0 else struct.unpack_from("<" + fmt, self.buf, self.pos + o)[0]

    def _target(self, i):
        o = self._off(i)
        if o == 0:
            return None
        p = self.pos + o
        return p + struct.unpack_from("<I", self.buf, p)[0]

    def table(self, i):
        t = self._target(i)
        return None if t is None else _Table(self.buf, t)

    def vector(self, i, dtype):
        t = self._target(i)
        if t is None:
            return None
        n = struct.unpack_from("<I", self.buf, t)[0]
        return np.frombuffer(self.buf, dtype=dtype, count=n, offset=t + 4).copy()


def parse_serialized(data):
    """Fields of a serialized quantizer (Quantizer / SphericalQuantizer / Transform tables):
    nbits, metric, centroid, mean_norm, pre_scale, kind and the transform's arrays."""
    buf = bytes(data)
    if buf[4:8] != b"sq00":
        raise ValueError("not a serialized spherical quantizer (identifier %r)" % buf[4:8])
    root = _Table(buf, struct.unpack_from("<I", buf, 0)[0])
    q = root.table(0)
    out = {"nbits": root.scalar(1, "I"), "centroid": q.vector(0, np.float32),
           "metric": _FB_METRIC[q.scalar(2, "b")], "mean_norm": q.scalar(3, "f", 0.0),
           "pre_scale": q.scalar(4, "f", 0.0), "signs0": None, "signs1": None,
           "padded_dim": 0, "subsample": None}
    tr = q.table(1)
    kind = _FB_TRANSFORM[tr.scalar(0, "B")]
    t = tr.table(1)
    out["kind"] = kind
    if kind == "double_hadamard":
        out["signs0"] = t.vector(0, np.uint8).astype(bool)
        out["signs1"] = t.vector(1, np.uint8).astype(bool)
        out["subsample"] = t.vector(2, np.uint32)
    elif kind == "padding_hadamard":
        out["signs0"] = t.vector(0, np.uint8).astype(bool)
        out["padded_dim"] = t.scalar(1, "I")
        out["subsample"] = t.vector(2, np.uint32)
    elif kind == "null":
        out["dim"] = t.scalar(0, "I")
    return out


class SphericalQuantizer:
    """1-bit spherical quantizer held on the device."""

    def __init__(self, metric, centroid, mean_norm, pre_scale=1.0, transform="double_hadamard",
                 signs0=None, signs1=None, padded_dim=0, subsample=None):
        self.lib = _native.load()
        self.metric = METRICS[metric] if isinstance(metric, str) else int(metric)
        self.centroid = np.ascontiguousarray(centroid, dtype=np.float32)
        self.dim = self.centroid.shape[0]
        self.transform = transform
        s0 = None if signs0 is None else np.ascontiguousarray(signs0, dtype=np.uint8)
        s1 = None if signs1 is None else np.ascontiguousarray(signs1, dtype=np.uint8)
        sub = None if subsample is None else np.ascontiguousarray(subsample, dtype=np.uint32)
        self.mean_norm, self.pre_scale = float(np.float32(mean_norm)), float(np.float32(pre_scale))
        self.signs0 = None if s0 is None else s0.astype(bool)
        self.signs1 = None if s1 is None else s1.astype(bool)
        self.padded_dim, self.subsample = int(padded_dim), sub
        self._h = C.c_void_p()
        check(self.lib.dkm_sph_create(self.metric, self.dim, _p(self.centroid), float(mean_norm),
                                      float(pre_scale), TRANSFORMS[transform], _p(s0), _p(s1),
                                      0 if s1 is None else s1.shape[0], int(padded_dim), _p(sub),
                                      0 if sub is None else sub.shape[0], C.byref(self._h)),
              self.lib)
        od, cb = C.c_int64(), C.c_int64()
        check(self.lib.dkm_sph_dims(self._h, C.byref(od), C.byref(cb)), self.lib)
        self.out_dim, self.code_bytes = od.value, cb.value

    @classmethod
    def from_serialized(cls, data):
        """A quantizer serialized by the reference (SphericalQuantizer::serialize)."""
        f = parse_serialized(data)
        if f["nbits"] != 1:
            raise ValueError("only 1-bit quantizers are supported (got %d bits)" % f["nbits"])
        if f["kind"] == "random_rotation":
            raise ValueError("the random-rotation transform is not supported")
        return cls(f["metric"], f["centroid"], f["mean_norm"], f["pre_scale"], f["kind"],
                   f["signs0"], f["signs1"], f["padded_dim"], f["subsample"])

    @classmethod
    def train(cls, data, metric, transform="double_hadamard", seed=0):
        """Centroid and mean norm as SphericalQuantizer::train computes them (on the
        device); the transform's random signs come from numpy's generator seeded with
        `seed`, not the reference's (its StdRng stream is not reproduced)."""
        lib = _native.load()
        x = np.ascontiguousarray(data, dtype=np.float32)
        m = METRICS[metric] if isinstance(metric, str) else int(metric)
        c = np.zeros(x.shape[1], np.float32)
        mn = C.c_float()
        check(lib.dkm_sph_train_stats(m, _p(x), x.shape[0], x.shape[1], _p(c), C.byref(mn)), lib)
        rng = np.random.default_rng(seed)
        dim = x.shape[1]
        if transform == "null":
            return cls(m, c, mn.value, 1.0, "null")
        if transform == "padding_hadamard":
            padded = 1 << (dim - 1).bit_length()
            return cls(m, c, mn.value, 1.0, transform, rng.integers(0, 2, dim), None, padded)
        return cls(m, c, mn.value, 1.0, transform, rng.integers(0, 2, dim),
                   rng.integers(0, 2, dim))

    def compress(self, x):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        out = np.zeros((x.shape[0], self.code_bytes), np.uint8)
        check(self.lib.dkm_sph_compress(self._h, _p(x), x.shape[0], _p(out)), self.lib)
        return out

    def distance(self, left, right, ids):
        left = np.ascontiguousarray(np.atleast_2d(left), dtype=np.uint8)
        right = np.ascontiguousarray(np.atleast_2d(right), dtype=np.uint8)
        ids = np.ascontiguousarray(np.atleast_2d(ids), dtype=np.uint32)
        out = np.zeros(ids.shape, np.float32)
        check(self.lib.dkm_sph_distance_batch(self._h, _p(left), left.shape[0], _p(right),
                                              right.shape[0], _p(ids), ids.shape[1], _p(out)),
              self.lib)
        return out

    def query_distance(self, queries, codes, ids, layout="four_bit_transposed",
                       allow_rescale=False):
        q = np.ascontiguousarray(np.atleast_2d(queries), dtype=np.float32)
        codes = np.ascontiguousarray(np.atleast_2d(codes), dtype=np.uint8)
        ids = np.ascontiguousarray(np.atleast_2d(ids), dtype=np.uint32)
        out = np.zeros(ids.shape, np.float32)
        check(self.lib.dkm_sph_query_distance_batch(self._h, LAYOUTS[layout], int(allow_rescale),
                                                    _p(q), q.shape[0], _p(codes), codes.shape[0],
                                                    _p(ids), ids.shape[1], _p(out)), self.lib)
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.dkm_sph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
