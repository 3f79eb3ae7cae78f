"""8-bit scalar quantization on the device (diskann-quantization/src/scalar).

    ScalarQuantizationParameters::train      scalar/train.rs:75-88
    ScalarQuantizer::compress_into           scalar/quantizer.rs:190-240, 380-432
    Compensated{SquaredL2, IP, CosineNormalized}   scalar/vectors.rs:182-470

A compressed vector is the reference's canonical layout: 4 bytes of f32 compensation,
then one code byte per dimension (MutCompensatedVectorRef::from_canonical_front_mut).
"""
import ctypes as C

import numpy as np

from . import _native
from .errors import check


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class ScalarQuantizer:
    """scale (dynamic range), shift[dim], mean_norm (query rescale for inner product)."""

    def __init__(self, scale, shift, mean_norm=None):
        self.scale = float(np.float32(scale))
        self.shift = np.ascontiguousarray(shift, dtype=np.float32)
        self.mean_norm = None if mean_norm is None else float(np.float32(mean_norm))

    @property
    def dim(self):
        return self.shift.shape[0]

    @classmethod
    def train(cls, data, standard_deviations=2.0):
        lib = _native.load()
        x = np.ascontiguousarray(data, dtype=np.float32)
        n, dim = x.shape
        shift = np.zeros(dim, np.float32)
        scale, mean_norm = C.c_float(), C.c_float()
        check(lib.dkm_sq_train(_p(x), n, dim, float(standard_deviations), C.byref(scale),
                               _p(shift), C.byref(mean_norm)), lib)
        return cls(scale.value, shift, mean_norm.value)

    def compress(self, x):
        """Canonical compressed rows, uint8[n, 4 + dim]."""
        lib = _native.load()
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        out = np.zeros((x.shape[0], 4 + self.dim), np.uint8)
        check(lib.dkm_sq_compress(self.scale, _p(self.shift), self.dim, _p(x), x.shape[0],
                                  _p(out)), lib)
        return out

    def distance(self, metric, left, right, ids):
        """out[i, j] = d(left[i], right[ids[i, j]]) for canonical compressed rows."""
        lib = _native.load()
        left = np.ascontiguousarray(np.atleast_2d(left), dtype=np.uint8)
        right = np.ascontiguousarray(np.atleast_2d(right), dtype=np.uint8)
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        if ids.ndim == 1:
            ids = ids[None, :]
        out = np.zeros(ids.shape, np.float32)
        check(lib.dkm_sq_distance_batch(int(metric), self.scale, _p(self.shift), self.dim,
                                        _p(left), left.shape[0], _p(right), right.shape[0],
                                        _p(ids), ids.shape[1], _p(out)), lib)
        return out


def split(compressed):
    """(codes uint8[n, dim], compensation float32[n]) of canonical compressed rows."""
    c = np.ascontiguousarray(np.atleast_2d(compressed), dtype=np.uint8)
    comp = c[:, :4].copy().view(np.float32).reshape(-1)
    return c[:, 4:].copy(), comp
