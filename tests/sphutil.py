"""Spherical-quantizer test helpers: a minimal flatbuffer reader for the reference's
serialized quantizer (diskann-quantization/schemas/spherical_v001.fbs and
transforms_v001.fbs) and the committed 1-bit golden baselines."""
import json
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "spherical_1bit.json")
METRICS = {"squared_l2": 2, "inner_product": 1, "cosine": 0}   # oracle metric numbers
FB_METRIC = {0: "squared_l2", 1: "inner_product", 2: "cosine"}
TRANSFORM_KIND = {1: "random_rotation", 2: "padding_hadamard", 3: "double_hadamard", 4: "null"}


class Table:
    def __init__(self, buf, pos):
        self.buf, self.pos = buf, pos
        vt = pos - struct.unpack_from("<i", buf, pos)[0]
        self.vt_len = struct.unpack_from("<H", buf, vt)[0]
        self.vt = vt

    def _off(self, i):
        o = 4 + 2 * i
        if o >= self.vt_len:
            return 0
        return struct.unpack_from("<H", self.buf, self.vt + o)[0]

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
        return None if t is None else Table(self.buf, t)

    def vector(self, i, dtype):
        t = self._target(i)
        if t is None:
            return None
        n = struct.unpack_from("<I", self.buf, t)[0]
        return np.frombuffer(self.buf, dtype=dtype, count=n, offset=t + 4).copy()


def parse_quantizer(data):
    """dict of the serialized SphericalQuantizer: nbits, metric, centroid, mean_norm,
    pre_scale, transform kind and its sign / subsample arrays."""
    buf = bytes(data)
    assert buf[4:8] == b"sq00"
    root = Table(buf, struct.unpack_from("<I", buf, 0)[0])
    q = root.table(0)
    out = {"nbits": root.scalar(1, "I"), "centroid": q.vector(0, np.float32),
           "metric": FB_METRIC[q.scalar(2, "b")], "mean_norm": q.scalar(3, "f", 0.0),
           "pre_scale": q.scalar(4, "f", 0.0)}
    tr = q.table(1)
    kind = TRANSFORM_KIND[tr.scalar(0, "B")]
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


def load():
    return json.load(open(GOLDEN))
