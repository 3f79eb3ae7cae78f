"""Candidate pre-scoring with exact FP32 rescore (dkm_prescore_*): exhaustive top-k where
every (query, row) pair is scored approximately in BF16 or FP8 (E4M3, power-of-two scaled
operands) on the matrix cores and only the best `candidates` rows of each query are scored
exactly with the reference's f32 distance functions (the arithmetic of Rerank,
full_precision.rs:357-404)."""
import ctypes as C

import numpy as np

from . import _native
from .errors import check
from .quantization import BF16, FP8, _Dev  # noqa: F401


class PreScorer:
    """Holds the rows (device copy) and their bf16 or fp8 copy.

    `metric`: Metric.L2, Metric.INNER_PRODUCT or Metric.COSINE_NORMALIZED; `precision`:
    quantization.BF16 or quantization.FP8."""

    def __init__(self, rows, metric, precision=BF16):
        self.lib = _native.load()
        x = np.ascontiguousarray(rows, dtype=np.float32)
        self.n, self.dim = x.shape
        self._x = _Dev(self.lib, x)
        h = C.c_void_p()
        check(self.lib.dkm_prescore_create(self._x.ptr, self.n, self.dim, int(metric),
                                           int(precision), C.byref(h)), self.lib)
        self._h = h

    def search(self, queries, k, candidates=4096):
        """Returns (ids[nq, k], distances[nq, k], stage_ms[4]); stage_ms holds the device
        time of the approximate scoring + filter, candidate selection, exact rescore and
        final selection."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q[None, :]
        if q.shape[1] != self.dim:
            raise ValueError("query dim %d != %d" % (q.shape[1], self.dim))
        nq = q.shape[0]
        dq = _Dev(self.lib, q)
        di = _Dev(self.lib, nbytes=4 * nq * k)
        dd = _Dev(self.lib, nbytes=4 * nq * k)
        ms = (C.c_double * 4)()
        check(self.lib.dkm_prescore_search(self._h, dq.ptr, nq, int(k), int(candidates), di.ptr,
                                           dd.ptr, ms), self.lib)
        return di.get(np.uint32, (nq, k)), dd.get(np.float32, (nq, k)), list(ms)

    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            self.lib.dkm_prescore_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
