"""Pin the oracle's disk-index search path: the transposed PQ table
(TransposedTable::process_into, diskann-quantization/src/product/tables/transposed/
table.rs:285-313, pivots.rs:785-820,915-948) and the DiskAccessor search with
RerankAndFilter (diskann-disk/src/search/provider/disk_provider.rs:323-388,666-721).

- Table entries: the fold acc = fma(p - q, p - q, acc) over a chunk's dimensions is
  checked against exact rational arithmetic rounded to f32 once per fma (fma is correctly
  rounded), on values where that order differs from the pairwise sum of the in-memory
  table (so a swapped order fails).
- Search: with one dimension per chunk both tables are the same numbers, so the disk
  traversal must take exactly the in-memory Hybrid traversal's cmps / hops from the same
  start row, and its results are that search's reranked results with the start row kept;
  with L above the row count every row is reranked, so the results are the brute-force
  top k; an io budget of 0 expands nothing (only the start row is reranked).
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle as po


def f32(x):
    return float(np.float32(x))


def rn_fma(a, b, c):
    """Correctly rounded f32 fma from exact rationals (ties to even through numpy's f64 ->
    f32 rounding of the exactly rounded f64 is not enough in general, so round via the
    nearest representable neighbours)."""
    exact = Fraction(a) * Fraction(b) + Fraction(c)
    lo = np.float32(float(exact))
    cands = [np.nextafter(lo, np.float32(-np.inf)), lo, np.nextafter(lo, np.float32(np.inf))]
    best = min(cands, key=lambda v: (abs(Fraction(float(v)) - exact),
                                     int(np.float32(v).view(np.uint32)) & 1))
    return float(best)


def test_transposed_table_is_a_fold_of_fmas():
    rng = np.random.default_rng(3)
    dim, M = 6, 3                     # chunk width 2
    piv = rng.standard_normal((256, dim)).astype(np.float32)
    q = rng.standard_normal(dim).astype(np.float32)
    off = po.chunk_offsets(dim, M)
    got = po.pq_lut_transposed(po.L2, piv, off, q).reshape(M, 256)
    inmem = po.pq_lut(po.L2, piv, off, q).reshape(M, 256)
    differs = 0
    for m in range(M):
        for c in range(256):
            acc = 0.0
            for d in range(off[m], off[m + 1]):
                t = f32(np.float32(piv[c, d]) - np.float32(q[d]))
                acc = rn_fma(t, t, acc)
            assert got[m, c] == np.float32(acc), (m, c)
            differs += got[m, c] != inmem[m, c]
    assert differs > 0          # the two tables' orders are distinguishable on this data
    ip = po.pq_lut_transposed(po.INNER_PRODUCT, piv, off, q).reshape(M, 256)
    for m in range(M):
        for c in range(0, 256, 17):
            acc = 0.0
            for d in range(off[m], off[m + 1]):
                acc = rn_fma(-float(q[d]), float(piv[c, d]), acc)
            assert ip[m, c] == np.float32(acc), (m, c)


def graph_index(n=1500, dim=24, seed=5):
    rng = np.random.default_rng(seed)
    cen = rng.standard_normal((8, dim)).astype(np.float32) * 3
    x = (cen[rng.integers(0, 8, n)] + rng.standard_normal((n, dim))).astype(np.float32)
    med, _ = po.medoid(x)
    vec = np.vstack([x, x[med:med + 1]])
    adj, deg = po.build_graph(po.L2, vec, R=16, R_max=20, L_build=32, alpha=1.2, starts=[n],
                              ids=np.arange(n), batch=64)
    qs = (cen[rng.integers(0, 8, 20)] + rng.standard_normal((20, dim))).astype(np.float32)
    return x, adj[:n], deg[:n], med, qs


def test_disk_traversal_equals_inmem_with_width_one_chunks():
    x, adj, deg, med, qs = graph_index()
    n, dim = x.shape
    off = po.chunk_offsets(dim, dim)           # one dimension per chunk
    rng = np.random.default_rng(1)
    piv = x[rng.choice(n, 256, replace=False)]
    codes = po.pq_encode(piv, off, x)
    for q in qs:
        assert np.array_equal(po.pq_lut_transposed(po.L2, piv, off, q),
                              po.pq_lut(po.L2, piv, off, q).reshape(-1))
    disk = po.OracleIndex(po.L2, x, adj, deg, [med], codes, piv, off)
    for i, q in enumerate(qs):
        for k, L, bw in ((10, 40, 1), (5, 20, 2)):
            ids, ds, cm, hp = disk.search_disk(q, k, L, bw)
            full = disk.search(q, L, L, bw, po.MODE_HYBRID)     # start row filtered
            assert (cm, hp) == (full[2], full[3]), i
            # the disk results are the Hybrid rerank with the start row put back
            d_med = po.distance(po.L2, q, x[med])
            merged = sorted(list(zip(full[1], full[0])) + [(d_med, med)])[:k]
            assert list(ids) == [int(b) for _, b in merged], i
            assert ds.tobytes() == np.array([a for a, _ in merged], np.float32).tobytes(), i


def test_disk_search_with_l_above_n_is_brute_force():
    x, adj, deg, med, qs = graph_index(n=300)
    off = po.chunk_offsets(x.shape[1], 8)
    rng = np.random.default_rng(2)
    piv = x[rng.choice(x.shape[0], 256, replace=False)]
    codes = po.pq_encode(piv, off, x)
    disk = po.OracleIndex(po.L2, x, adj, deg, [med], codes, piv, off)
    for q in qs[:5]:
        ids, ds, _, _ = disk.search_disk(q, 10, 400, 1)
        exact = np.array([po.distance(po.L2, q, v) for v in x], np.float32)
        order = np.argsort(exact, kind="stable")[:10]
        assert list(ids) == list(order)
        assert ds.tobytes() == exact[order].tobytes()


def test_disk_search_io_budget():
    x, adj, deg, med, qs = graph_index(n=400)
    off = po.chunk_offsets(x.shape[1], 8)
    rng = np.random.default_rng(4)
    piv = x[rng.choice(x.shape[0], 256, replace=False)]
    codes = po.pq_encode(piv, off, x)
    disk = po.OracleIndex(po.L2, x, adj, deg, [med], codes, piv, off)
    ids, ds, cm, hp = disk.search_disk(qs[0], 5, 20, 1, io_limit=0)
    assert (list(ids), cm, hp) == ([med], 1, 1)
    assert ds[0] == po.distance(po.L2, qs[0], x[med])
    # a budget of b loads expands exactly b beam nodes; later beam nodes count as hops
    free = disk.search_disk(qs[1], 5, 30, 2)
    for b in (1, 3, 7):
        _, _, cm_b, hp_b = disk.search_disk(qs[1], 5, 30, 2, io_limit=b)
        assert cm_b <= free[2] and hp_b >= b
    with pytest.raises(ValueError):
        disk.search_disk(qs[0], 10, 5, 1)             # L < k


def test_closest_centers_exact_on_integer_data():
    # small integers: every product, sum and norm is exact in f32, so the form
    # (|x|^2 + |c|^2) - 2<x,c> is the exact squared distance and the result is the
    # brute-force order (ties by lower centre index), for k = 1 and k > 1
    rng = np.random.default_rng(8)
    x = rng.integers(-6, 7, (500, 20)).astype(np.float32)
    c = rng.integers(-6, 7, (40, 20)).astype(np.float32)
    c[7] = c[3]                      # duplicate centres: the lower index first
    exact = ((x[:, None, :].astype(np.int64) - c[None, :, :].astype(np.int64)) ** 2).sum(-1)
    order = np.argsort(exact, axis=1, kind="stable")
    assert np.array_equal(po.closest_centers(x, c, 1)[:, 0], order[:, 0])
    assert np.array_equal(po.closest_centers(x, c, 5), order[:, :5])


def test_closest_centers_float_data_within_rounding():
    # float data: the k = 1 answer is the FP64 argmin wherever the FP64 margin to the
    # runner-up exceeds the f32 error of the three-term form
    rng = np.random.default_rng(9)
    x = rng.standard_normal((400, 64)).astype(np.float32)
    c = rng.standard_normal((32, 64)).astype(np.float32)
    got = po.closest_centers(x, c, 1)[:, 0]
    d64 = ((x[:, None, :].astype(np.float64) - c[None, :, :]) ** 2).sum(-1)
    srt = np.sort(d64, axis=1)
    bound = 1e-5 * (np.abs(x).sum(1) ** 2 + (np.abs(c).sum(1) ** 2).max())
    clear = (srt[:, 1] - srt[:, 0]) > bound
    assert clear.mean() > 0.9
    assert np.array_equal(got[clear], np.argmin(d64, axis=1)[clear])
