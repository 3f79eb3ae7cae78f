"""Batched search captured in a HIP graph (dkm_search_graph_*): every replay returns exactly
what the eager batched search returns for the same queries (ids, f32 distances, counts,
cmps, hops) and the oracle agrees; replays are interleaved with eager launches and host
synchronisation and see fresh query contents; a replay after the index's launch changed
is refused."""
import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd.errors import ANNError
from oracle import pyoracle as po
from test_gpu_search import dataset, log, make_index, pq_setup

pytestmark = pytest.mark.gpu


def same(a, b):
    for i in range(a.ids.shape[0]):
        c = int(a.counts[i])
        assert c == int(b.counts[i]), i
        assert np.array_equal(a.ids[i, :c], b.ids[i, :c]), i
        assert a.distances[i, :c].tobytes() == b.distances[i, :c].tobytes(), i
    assert np.array_equal(a.cmps, b.cmps) and np.array_equal(a.hops, b.hops)


@pytest.mark.parametrize("strategy", [dk.FullPrecision, dk.Hybrid])
def test_graph_replays_equal_eager_and_oracle(strategy):
    log("search graph strategy=%d" % strategy)
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    piv, off = pq_setup(vec, 32)
    ix.set_pq(piv, off)
    ix.encode()
    oix = po.OracleIndex(po.L2, vec, adj, deg, [vec.shape[0] - 1], ix.get_codes(), piv, off)
    mode = po.MODE_FULL if strategy == dk.FullPrecision else po.MODE_HYBRID
    rng = np.random.default_rng(5)
    nq = 48
    ks = rng.choice([1, 5, 10], nq).astype(np.uint32)
    Ls = np.maximum(ks, rng.integers(8, 96, nq)).astype(np.uint32)
    batches = [qs[:nq], qs[-nq:][::-1].copy(),
               rng.standard_normal((nq, vec.shape[1])).astype(np.float32)]
    with dk.SearchGraph(ix, batches[0], ks, Ls, 2, strategy, max_k=10, max_L=96) as g:
        for rep, q in enumerate(batches + batches[:1]):
            g.set_queries(q)
            got = g.run()
            # eager work on the same stream between replays
            want = ix.search_batch(q, ks, Ls, 2, strategy)
            same(got, want)
            for i in (0, nq // 2, nq - 1):
                ids, ds, cm, hp = oix.search(q[i], int(ks[i]), int(Ls[i]), 2, mode)
                assert list(got.ids[i, :got.counts[i]]) == list(ids), (rep, i)
                assert (int(got.cmps[i]), int(got.hops[i])) == (cm, hp), (rep, i)


def test_graph_refuses_a_changed_index():
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    ix.set_tuning(512, 1)
    with dk.SearchGraph(ix, qs[:8], 5, 32, 1, dk.FullPrecision) as g:
        g.run()
        ix.set_tuning(1024, 1)  # a different launch shape: the capture is stale
        with pytest.raises(ANNError):
            g.run()
        ix.set_tuning(512, 1)
        g.run()


def test_graph_needs_queries():
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    with pytest.raises(ANNError):
        dk.SearchGraph(ix, qs[:0], 5, 32, 1, dk.FullPrecision)
    with pytest.raises(dk.KnnSearchError):
        dk.SearchGraph(ix, qs[:4], 5, 4, 1, dk.FullPrecision)  # L < k (Knn::new)
