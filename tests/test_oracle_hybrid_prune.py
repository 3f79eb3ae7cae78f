"""Hybrid prune on the oracle (HybridPruneAccessor, inmem/product.rs:144-221): the first
max_fp ids a prune fills are full precision, the rest are compared through the direct PQ
distances (HybridComputer, async_/distances.rs:118-165).

Pins:
* max_fp at least the pool size: every id full precision, so the graph equals the full-
  precision build exactly.
* Data whose rows are exactly the vectors their PQ codes reconstruct, on an integer pivot
  table: every direct / code-vs-code distance equals the full-precision one, so the
  graph is the same for every max_fp (0, 2, 32) - a flag or operand mix-up breaks this.
* Real-valued data with max_fp = 2: a valid graph (no self loops or duplicates, degree
  <= R_max) that differs from the full-precision one (the PQ distances are used)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from test_oracle_pq_direct import expected_vector, seed_pivot_table


def build(vectors, max_fp=None, codes=None, pivots=None, off=None, batch=16, metric=po.L2):
    n = vectors.shape[0] - 1
    return po.build_graph(metric, vectors, 12, 16, 30, 1.2, np.array([n], np.uint32),
                          np.arange(n, dtype=np.uint32), batch=batch, intra_batch=-1,
                          start_adj={n: []}, max_fp=max_fp, codes=codes, pivots=pivots,
                          pq_offsets=off)


def reconstructed_dataset(n, dim, M, K, seed):
    piv, off = seed_pivot_table(dim, M, K)
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, K, (n + 1, M)).astype(np.uint8)
    x = np.stack([expected_vector(c, off) for c in codes]).astype(np.float32)
    return x, codes, piv, off


@pytest.mark.parametrize("batch", [1, 16])
def test_all_full_precision_equals_fp_build(batch):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((300, 16)).astype(np.float32)
    piv = rng.standard_normal((256, 16)).astype(np.float32)
    off = po.chunk_offsets(16, 4)
    codes = po.pq_encode(piv, off, x)
    a0, d0 = build(x, batch=batch)
    a1, d1 = build(x, 100000, codes, piv, off, batch=batch)
    assert np.array_equal(d0, d1) and np.array_equal(a0, a1)


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT])
@pytest.mark.parametrize("max_fp", [0, 2, 32])
def test_exact_codes_give_fp_graph(metric, max_fp):
    x, codes, piv, off = reconstructed_dataset(240, 12, 4, 20, 7)
    a0, d0 = build(x, metric=metric)
    a1, d1 = build(x, max_fp, codes, piv, off, metric=metric)
    assert np.array_equal(d0, d1) and np.array_equal(a0, a1)


def test_real_valued_hybrid_graph_is_valid_and_differs():
    rng = np.random.default_rng(3)
    n, dim = 400, 16
    x = rng.standard_normal((n + 1, dim)).astype(np.float32)
    piv = rng.standard_normal((16, dim)).astype(np.float32)
    off = po.chunk_offsets(dim, 4)
    codes = po.pq_encode(piv, off, x)
    a0, d0 = build(x)
    a1, d1 = build(x, 2, codes, piv, off)
    for i in range(n + 1):
        lst = a1[i, :d1[i]].tolist()
        assert i not in lst and len(set(lst)) == len(lst) and len(lst) <= 16
    assert not (np.array_equal(d0, d1) and np.array_equal(a0, a1))
