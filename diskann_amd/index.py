"""Host-side mirror of the reference's search interface, driving the HIP kernels.

    Knn::new(k, L, beam_width)          diskann/src/graph/search/knn_search.rs:94-130
    DiskANNIndex::search / search_with  diskann/src/graph/index.rs:2034-2070
    FullPrecision / Hybrid / Quantized  diskann-providers/src/model/graph/provider/async_/inmem/
    DefaultProvider storage             .../inmem/provider.rs:227-248 (frozen start rows after
                                        the data rows, set_start_points :393-421)
"""
import ctypes as C
import enum

import numpy as np

from . import _native
from .errors import (BEAM_WIDTH_ZERO, K_ZERO, L_LESS_THAN_K, L_ZERO, KnnSearchError, check)
from .quantization import _Dev


class Metric(enum.IntEnum):
    """diskann-vector/src/distance/metric.rs:11-20 (repr(C) order)."""
    COSINE = 0
    INNER_PRODUCT = 1
    L2 = 2
    COSINE_NORMALIZED = 3


class Strategy(enum.IntEnum):
    FULL_PRECISION = 0   # FullPrecision: FP32 traversal, FilterStartPoints + CopyIds
    HYBRID = 1           # Hybrid: PQ traversal, FilterStartPoints + Rerank (exact FP32)
    QUANTIZED = 2        # Quantized: PQ traversal, FilterStartPoints + copy PQ distances


FullPrecision = Strategy.FULL_PRECISION
Hybrid = Strategy.HYBRID
Quantized = Strategy.QUANTIZED


class Knn:
    """k-NN search parameters with the reference's validation and error kinds."""

    def __init__(self, k, L, beam_width=None):
        if k == 0:
            raise KnnSearchError(K_ZERO, "k must be non-zero")
        if L == 0:
            raise KnnSearchError(L_ZERO, "L must be non-zero")
        if k > L:
            raise KnnSearchError(L_LESS_THAN_K, "L (%d) is less than k (%d)" % (L, k))
        bw = 1 if beam_width is None else beam_width
        if bw == 0:
            raise KnnSearchError(BEAM_WIDTH_ZERO, "beam width must be non-zero")
        self.k, self.L, self.beam_width = int(k), int(L), int(bw)

    def __repr__(self):
        return "Knn(k=%d, L=%d, beam_width=%d)" % (self.k, self.L, self.beam_width)


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class SearchResult:
    def __init__(self, ids, distances, counts, cmps, hops):
        self.ids, self.distances, self.counts, self.cmps, self.hops = (
            ids, distances, counts, cmps, hops)


class DiskANNIndex:
    """Device-resident in-memory index: N data rows + `num_start` frozen start rows
    (ids N, N+1, ...), fixed-stride adjacency, optional PQ schema and codes."""

    def __init__(self, metric, dim, max_points, R, R_max=None, num_start=1, device=0):
        self.lib = _native.load()
        check(self.lib.dkm_init(device), self.lib)
        self.metric = Metric(metric)
        self.dim, self.max_points, self.num_start = int(dim), int(max_points), int(num_start)
        self.R = int(R)
        self.R_max = int(R_max if R_max is not None else int(1.3 * R))
        h = C.c_void_p()
        check(self.lib.dkm_index_create(int(self.metric), self.dim, self.max_points,
                                        self.num_start, self.R, self.R_max, C.byref(h)),
              self.lib)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.dkm_index_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def start_ids(self):
        return np.arange(self.max_points, self.max_points + self.num_start, dtype=np.uint32)

    def set_vectors(self, x, row0=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != self.dim:
            raise ValueError("expected (n, %d) float32 vectors" % self.dim)
        check(self.lib.dkm_set_vectors(self._h, _p(x), row0, x.shape[0]), self.lib)

    def set_vectors_from_bin(self, path):
        """Streams an f32 .bin file into the data rows."""
        check(self.lib.dkm_set_vectors_from_bin(self._h, str(path).encode()), self.lib)

    def save_graph(self, path, start_point=None):
        """The in-memory layout (every row, start points last; save_direct)."""
        sp = self.max_points if start_point is None else int(start_point)
        check(self.lib.dkm_save_graph(self._h, str(path).encode(), sp), self.lib)

    def save_graph_disk(self, path, actual_start_point):
        """The disk-index layout (data rows only, start id remapped; DiskAdaptor)."""
        check(self.lib.dkm_save_graph_disk(self._h, str(path).encode(), int(actual_start_point)),
              self.lib)

    def load_graph(self, path):
        """Loads a graph file in the in-memory layout; returns the header's start id."""
        sp = C.c_uint32()
        check(self.lib.dkm_load_graph(self._h, str(path).encode(), C.byref(sp)), self.lib)
        return sp.value

    def set_start_points(self, vectors):
        v = np.ascontiguousarray(np.asarray(vectors, dtype=np.float32).reshape(self.num_start,
                                                                               self.dim))
        check(self.lib.dkm_set_start_points(self._h, _p(v)), self.lib)

    def set_graph(self, lists, counts, row0=0):
        lists = np.ascontiguousarray(lists, dtype=np.uint32)
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        check(self.lib.dkm_set_graph(self._h, _p(lists), _p(counts), lists.shape[1], row0,
                                     lists.shape[0]), self.lib)

    def get_graph(self, row0=0, n=None):
        n = self.max_points + self.num_start - row0 if n is None else n
        lists = np.zeros((n, self.R_max), dtype=np.uint32)
        counts = np.zeros(n, dtype=np.uint32)
        check(self.lib.dkm_get_graph(self._h, _p(lists), _p(counts), self.R_max, row0, n),
              self.lib)
        return lists, counts

    def set_pq(self, pivots, offsets):
        pivots = np.ascontiguousarray(pivots, dtype=np.float32)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        check(self.lib.dkm_set_pq(self._h, _p(pivots), pivots.shape[0], _p(offsets),
                                  offsets.shape[0] - 1), self.lib)
        self.M = offsets.shape[0] - 1

    def set_sq(self, quantizer):
        """Scalar-quantized store (SQStore, diskann-providers/.../inmem/scalar.rs:83-253):
        from now on the Hybrid / Quantized strategies traverse with 8-bit codes (call
        encode_sq or set_sq_codes next); use_pq switches back to the PQ store."""
        mn = 0.0 if quantizer.mean_norm is None else quantizer.mean_norm
        check(self.lib.dkm_set_sq(self._h, quantizer.scale, _p(quantizer.shift), mn), self.lib)

    def encode_sq(self):
        check(self.lib.dkm_encode_index_sq(self._h), self.lib)

    def set_sq_codes(self, compressed, row0=0):
        c = np.ascontiguousarray(np.atleast_2d(compressed), dtype=np.uint8)
        check(self.lib.dkm_set_sq_codes(self._h, _p(c), row0, c.shape[0]), self.lib)

    def get_sq_codes(self, row0=0, n=None):
        n = self.max_points + self.num_start - row0 if n is None else n
        out = np.zeros((n, 4 + self.dim), dtype=np.uint8)
        check(self.lib.dkm_get_sq_codes(self._h, _p(out), row0, n), self.lib)
        return out

    def set_spherical(self, quantizer, layout="four_bit_transposed"):
        """1-bit spherical store (SphericalStore, diskann-providers/.../inmem/spherical.rs):
        the Hybrid / Quantized strategies traverse with 1-bit codes and the query layout
        `same_as_data` or `four_bit_transposed` (call encode_spherical or set_sph_codes
        next)."""
        from .spherical import LAYOUTS
        check(self.lib.dkm_set_spherical(self._h, quantizer._h, LAYOUTS[layout]), self.lib)
        self._sph_bytes = quantizer.code_bytes

    def encode_spherical(self):
        check(self.lib.dkm_encode_index_sph(self._h), self.lib)

    def set_sph_codes(self, compressed, row0=0):
        c = np.ascontiguousarray(np.atleast_2d(compressed), dtype=np.uint8)
        check(self.lib.dkm_set_sph_codes(self._h, _p(c), row0, c.shape[0]), self.lib)

    def get_sph_codes(self, row0=0, n=None):
        n = self.max_points + self.num_start - row0 if n is None else n
        out = np.zeros((n, self._sph_bytes), dtype=np.uint8)
        check(self.lib.dkm_get_sph_codes(self._h, _p(out), row0, n), self.lib)
        return out

    def use_pq(self):
        check(self.lib.dkm_use_pq(self._h), self.lib)

    def set_codes(self, codes, row0=0):
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        check(self.lib.dkm_set_codes(self._h, _p(codes), row0, codes.shape[0]), self.lib)

    def encode(self):
        check(self.lib.dkm_encode_index(self._h), self.lib)

    def get_codes(self, row0=0, n=None):
        n = self.max_points + self.num_start - row0 if n is None else n
        out = np.zeros((n, self.M), dtype=np.uint8)
        check(self.lib.dkm_get_codes(self._h, _p(out), row0, n), self.lib)
        return out

    def set_build_hybrid(self, enabled=True, max_fp_vecs_per_prune=None):
        """Hybrid prune for build() (common::Hybrid::new(max_fp_vecs_per_prune)): each prune
        keeps the first max_fp_vecs_per_prune ids of its fill order in full precision and
        compares the others through their PQ codes; None = Hybrid::new(None). enabled=False
        restores full precision everywhere. Needs set_pq and the codes of every row."""
        m = -1 if not enabled else (2**31 - 1 if max_fp_vecs_per_prune is None
                                    else int(max_fp_vecs_per_prune))
        if enabled and m < 0:
            raise ValueError("max_fp_vecs_per_prune must be >= 0")
        check(self.lib.dkm_set_build_hybrid(self._h, m), self.lib)

    def set_build_sq(self, enabled=True):
        """Scalar-quantized build (the Quantized strategy over an SQStore,
        diskann-providers/src/model/graph/provider/async_/inmem/scalar.rs:676-720): the
        insert search traverses with the SQ query computer and every prune distance is the
        compensated distance between stored codes. Needs set_sq and encode_sq (codes of
        every row); the reference runs it with build(batch=1)."""
        check(self.lib.dkm_set_build_sq(self._h, int(bool(enabled))), self.lib)

    def set_build_spherical(self, enabled=True):
        """Spherical build (Quantized::build() over a SphericalStore,
        diskann-providers/src/model/graph/provider/async_/inmem/spherical.rs:515-533,
        605-650): the insert search traverses with SameAsData queries and no inner-product
        rescale, every prune distance is the data x data distance of stored codes. Needs
        set_spherical and encode_spherical."""
        check(self.lib.dkm_set_build_spherical(self._h, int(bool(enabled))), self.lib)

    def build(self, ids=None, batch=1, intra_batch=-1, L_build=100, alpha=1.2,
              max_occlusion=750, saturate=False):
        """Insert rows `ids` (default: every data row, in order) with multi_insert
        semantics in batches of `batch` (batch 1 = insert()); intra_batch -1 = All,
        0 = None, n = Max(n)."""
        ids = np.arange(self.max_points, dtype=np.uint32) if ids is None else \
            np.ascontiguousarray(ids, dtype=np.uint32)
        check(self.lib.dkm_build(self._h, _p(ids), ids.shape[0], int(batch), int(intra_batch),
                                 int(L_build), float(alpha), int(max_occlusion), int(saturate)),
              self.lib)

    def prune_batch(self, locations, lists, alpha=1.2, max_occlusion=750, saturate=False,
                    force_saturate=False):
        """robust_prune_list of each candidate list around its location (row id): returns
        a list of pruned id arrays. `lists` is a sequence of id arrays (one per location)."""
        loc = np.ascontiguousarray(locations, dtype=np.uint32).reshape(-1)
        if len(lists) != loc.shape[0]:
            raise ValueError("one candidate list per location")
        lens = np.array([len(x) for x in lists], dtype=np.uint32)
        stride = max(1, int(lens.max()) if lens.size else 1)
        buf = np.zeros((loc.shape[0], stride), dtype=np.uint32)
        for i, x in enumerate(lists):
            buf[i, :len(x)] = x
        out = np.zeros((loc.shape[0], max(self.R, 1)), dtype=np.uint32)
        cnt = np.zeros(loc.shape[0], dtype=np.uint32)
        check(self.lib.dkm_prune_batch(self._h, _p(loc), loc.shape[0], _p(buf), _p(lens), stride,
                                       float(alpha), int(max_occlusion), int(saturate),
                                       int(force_saturate), _p(out), _p(cnt), out.shape[1]),
              self.lib)
        return [out[i, :cnt[i]].copy() for i in range(loc.shape[0])]

    def prune_range(self, ids=None, alpha=1.2, max_occlusion=750, saturate=False):
        """DiskANNIndex::prune_range: prune every list of `ids` longer than R in place."""
        ids = np.arange(self.max_points, dtype=np.uint32) if ids is None else \
            np.ascontiguousarray(ids, dtype=np.uint32)
        check(self.lib.dkm_prune_range(self._h, _p(ids), ids.shape[0], float(alpha),
                                       int(max_occlusion), int(saturate)), self.lib)

    def medoid(self):
        """(row, vector) of compute_medoid over the data rows."""
        v = np.zeros(self.dim, dtype=np.float32)
        row = C.c_int64()
        check(self.lib.dkm_medoid(self._h, _p(v), C.byref(row)), self.lib)
        return int(row.value), v

    def bruteforce(self, queries, k):
        """Exact FP64 top-k over the data rows: (ids[nq, k], dists[nq, k] float64)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        ids = np.zeros((q.shape[0], k), dtype=np.uint32)
        ds = np.zeros((q.shape[0], k), dtype=np.float64)
        check(self.lib.dkm_bruteforce(self._h, _p(q), q.shape[0], int(k), _p(ids), _p(ds)),
              self.lib)
        return ids, ds

    def autotune(self, queries, k, L, beam_width=1, strategy=Strategy.HYBRID):
        """Time the search of `queries` under the result-preserving knobs and keep the
        fastest; returns (visited_lds_slots, adc_mode)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        d = C.c_void_p()
        check(self.lib.dkm_device_alloc(q.nbytes, C.byref(d)), self.lib)
        try:
            check(self.lib.dkm_memcpy_h2d(d, _p(q), q.nbytes), self.lib)
            hs = C.c_int64()
            am = C.c_int()
            check(self.lib.dkm_autotune(self._h, int(strategy), d, q.shape[0], int(k), int(L),
                                        int(beam_width), C.byref(hs), C.byref(am)), self.lib)
        finally:
            self.lib.dkm_device_free(d)
        self.tuning = (int(hs.value), int(am.value))
        return self.tuning

    def set_tuning(self, visited_lds_slots=0, adc_mode=1):
        """Performance knobs that never change results: LDS visited-set slots per query
        (0 = automatic) and PQ distance mode (0 = LDS lookup table, 1 = direct)."""
        self.tuning = (int(visited_lds_slots), int(adc_mode))
        check(self.lib.dkm_set_tuning(self._h, int(visited_lds_slots), int(adc_mode)), self.lib)

    def search_batch(self, queries, k, L, beam_width=1, strategy=Strategy.FULL_PRECISION):
        """Batched search. `k` and `L` are ints or per-query arrays (validated per query
        exactly like Knn::new)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q[None, :]
        if q.shape[1] != self.dim:
            raise ValueError("query dim %d != index dim %d" % (q.shape[1], self.dim))
        nq = q.shape[0]
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=np.uint32), (nq,)))
        Ls = np.ascontiguousarray(np.broadcast_to(np.asarray(L, dtype=np.uint32), (nq,)))
        kstride = int(ks.max()) if nq else 1
        ids = np.zeros((nq, kstride), dtype=np.uint32)
        ds = np.zeros((nq, kstride), dtype=np.float32)
        cnt = np.zeros(nq, dtype=np.uint32)
        cmps = np.zeros(nq, dtype=np.uint32)
        hops = np.zeros(nq, dtype=np.uint32)
        check(self.lib.dkm_search_batch(self._h, int(strategy), _p(q), nq, _p(ks), _p(Ls),
                                        int(beam_width), kstride, _p(ids), _p(ds), _p(cnt),
                                        _p(cmps), _p(hops)), self.lib)
        return SearchResult(ids, ds, cnt, cmps, hops)

    def disk_search_batch(self, queries, k, L, beam_width=1, medoid=0, io_limit=None):
        """Disk-index search (DiskIndexSearcher: PQ traversal from the data row `medoid` with
        the transposed table's arithmetic, at most io_limit expanded nodes per query, every
        queue entry reranked at full precision). k and L as in search_batch."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q[None, :]
        if q.shape[1] != self.dim:
            raise ValueError("query dim %d != index dim %d" % (q.shape[1], self.dim))
        nq = q.shape[0]
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=np.uint32), (nq,)))
        Ls = np.ascontiguousarray(np.broadcast_to(np.asarray(L, dtype=np.uint32), (nq,)))
        kstride = int(ks.max()) if nq else 1
        ids = np.zeros((nq, kstride), dtype=np.uint32)
        ds = np.zeros((nq, kstride), dtype=np.float32)
        cnt = np.zeros(nq, dtype=np.uint32)
        cmps = np.zeros(nq, dtype=np.uint32)
        hops = np.zeros(nq, dtype=np.uint32)
        lim = (1 << 64) - 1 if io_limit is None else int(io_limit)
        check(self.lib.dkm_disk_search_batch(self._h, _p(q), nq, _p(ks), _p(Ls), int(beam_width),
                                             int(medoid), lim, kstride, _p(ids), _p(ds), _p(cnt),
                                             _p(cmps), _p(hops)), self.lib)
        return SearchResult(ids, ds, cnt, cmps, hops)

    def search(self, knn, strategy, query):
        """Single-query form of DiskANNIndex::search(Knn, strategy, query)."""
        r = self.search_batch(query, knn.k, knn.L, knn.beam_width, strategy)
        n = int(r.counts[0])
        return r.ids[0, :n], r.distances[0, :n], int(r.cmps[0]), int(r.hops[0])


class SearchGraph:
    """A batched search captured in a HIP graph (dkm_search_graph_*): the queries (n x dim
    float32), k and L of each query are copied into device buffers owned by the graph;
    `set_queries` refreshes them and `run` replays the captured launch and returns the same
    SearchResult as DiskANNIndex.search_batch for those queries."""

    def __init__(self, index, queries, k, L, beam_width=1, strategy=Strategy.FULL_PRECISION,
                 max_k=None, max_L=None):
        self.index, self.lib = index, index.lib
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim != 2 or q.shape[1] != index.dim:
            raise ValueError("queries must be (n, %d)" % index.dim)
        self.nq = q.shape[0]
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=np.uint32), (self.nq,)))
        Ls = np.ascontiguousarray(np.broadcast_to(np.asarray(L, dtype=np.uint32), (self.nq,)))
        for kk, ll in set(zip(ks.tolist(), Ls.tolist())):
            Knn(kk, ll, beam_width)  # Knn::new validation (the device path does none)
        self.max_k = (int(ks.max()) if self.nq else 1) if max_k is None else int(max_k)
        self.max_L = (int(Ls.max()) if self.nq else 1) if max_L is None else int(max_L)
        if self.nq and (self.max_k < int(ks.max()) or self.max_L < int(Ls.max())):
            raise ValueError("max_k / max_L below a query's k / L")
        self._dq = _Dev(self.lib, q)
        self._dk = _Dev(self.lib, ks)
        self._dL = _Dev(self.lib, Ls)
        n = self.nq
        self._ids = _Dev(self.lib, nbytes=4 * n * self.max_k)
        self._d = _Dev(self.lib, nbytes=4 * n * self.max_k)
        self._cnt = _Dev(self.lib, nbytes=4 * n)
        self._cm = _Dev(self.lib, nbytes=4 * n)
        self._hp = _Dev(self.lib, nbytes=4 * n)
        h = C.c_void_p()
        check(self.lib.dkm_search_graph_create(
            index._h, int(strategy), self._dq.ptr, n, self._dk.ptr, self._dL.ptr, self.max_k,
            self.max_L, int(beam_width), self.max_k, self._ids.ptr, self._d.ptr, self._cnt.ptr,
            self._cm.ptr, self._hp.ptr, None, C.byref(h)), self.lib)
        self._h = h

    def set_queries(self, queries):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.shape != (self.nq, self.index.dim):
            raise ValueError("queries must be (%d, %d)" % (self.nq, self.index.dim))
        self._dq.put(q)

    def run(self):
        check(self.lib.dkm_search_graph_launch(self._h), self.lib)
        check(self.lib.dkm_search_graph_status(self._h, None), self.lib)
        n, mk = self.nq, self.max_k
        return SearchResult(self._ids.get(np.uint32, (n, mk)), self._d.get(np.float32, (n, mk)),
                            self._cnt.get(np.uint32, n), self._cm.get(np.uint32, n),
                            self._hp.get(np.uint32, n))

    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            self.lib.dkm_search_graph_destroy(h)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous batching over one index.

    Requests go into a ring of `capacity` entries in pinned host memory; every submit and
    poll launches the requests waiting at that moment as one batch on one of eight
    streams, while at most `slots` requests are in flight (0: twice what the device holds
    at once), so requests arriving while others run overlap with them on the device.
    Each result equals what `search_batch` returns for that query alone (same k, L, beam
    width and strategy). Submitting to a full ring raises CapacityError; `poll` returns
    finished requests. Progress needs the caller to keep calling submit or poll.
    """

    def __init__(self, index, strategy=Strategy.HYBRID, max_k=10, max_L=512, beam_width=1,
                 capacity=4096, slots=0):
        self.index, self.lib = index, index.lib
        self.dim, self.max_k = index.dim, int(max_k)
        h = C.c_void_p()
        check(self.lib.dkm_server_start(index._h, int(strategy), int(max_k), int(max_L),
                                        int(beam_width), int(capacity), int(slots), C.byref(h)),
              self.lib)
        self._h = h
        n = C.c_int64()
        check(self.lib.dkm_server_slots(h, C.byref(n)), self.lib)
        self.slots = n.value

    def submit(self, query, k, L, tag):
        q = np.ascontiguousarray(query, dtype=np.float32).reshape(-1)
        if q.shape[0] != self.dim:
            raise ValueError("query dim %d != index dim %d" % (q.shape[0], self.dim))
        check(self.lib.dkm_server_submit(self._h, _p(q), int(k), int(L), int(tag)), self.lib)

    def submit_batch(self, queries, k, L, tags):
        """Submits in order until the ring is full; returns how many were accepted."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim != 2 or q.shape[1] != self.dim:
            raise ValueError("queries must be (n, %d)" % self.dim)
        n = q.shape[0]
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=np.uint32), (n,)))
        Ls = np.ascontiguousarray(np.broadcast_to(np.asarray(L, dtype=np.uint32), (n,)))
        tg = np.ascontiguousarray(np.broadcast_to(np.asarray(tags, dtype=np.uint64), (n,)))
        acc = C.c_int64()
        check(self.lib.dkm_server_submit_batch(self._h, _p(q), n, _p(ks), _p(Ls), _p(tg),
                                               C.byref(acc)), self.lib)
        return acc.value

    def poll(self, max_results=4096):
        """Finished requests: (tags, ids[n, max_k], dists[n, max_k], counts, cmps, hops,
        latency_us); latency is measured from submit to the poll that returned it."""
        m = int(max_results)
        if getattr(self, "_pbuf", None) is None or self._pbuf[0].shape[0] < m:
            self._pbuf = (np.empty(m, np.uint64), np.empty((m, self.max_k), np.uint32),
                          np.empty((m, self.max_k), np.float32), np.empty(m, np.uint32),
                          np.empty(m, np.uint32), np.empty(m, np.uint32),
                          np.empty(m, np.float64))
        tags, ids, ds, cnt, cmps, hops, lat = self._pbuf
        n = C.c_int64()
        check(self.lib.dkm_server_poll(self._h, m, _p(tags), _p(ids), _p(ds), _p(cnt), _p(cmps),
                                       _p(hops), _p(lat), C.byref(n)), self.lib)
        k = n.value
        return tuple(a[:k].copy() for a in self._pbuf)

    def pending(self):
        n = C.c_int64()
        check(self.lib.dkm_server_pending(self._h, C.byref(n)), self.lib)
        return n.value

    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            check(self.lib.dkm_server_stop(h), self.lib)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
