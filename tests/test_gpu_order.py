"""Launches of 2048 queries or more run them longest L first (block b takes query order[b],
a counting sort over L on the device). Each query's result must not depend on that: a
large mixed-L batch equals the same queries searched in small launches (which keep the
plain order), eagerly and as a HIP-graph replay."""
import numpy as np
import pytest

import diskann_amd as dk
from test_gpu_search import dataset, log, make_index, pq_setup
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def mixed(n, seed):
    rng = np.random.default_rng(seed)
    ks = rng.choice([1, 5, 10], n).astype(np.uint32)
    Ls = np.maximum(ks, rng.integers(8, 200, n)).astype(np.uint32)
    return ks, Ls


def check_same(big, parts, n):
    for name in ("counts", "cmps", "hops"):
        assert np.array_equal(getattr(big, name), getattr(parts, name)), name
    for i in range(n):
        c = int(big.counts[i])
        assert np.array_equal(big.ids[i, :c], parts.ids[i, :c]), i
        assert big.distances[i, :c].tobytes() == parts.distances[i, :c].tobytes(), i


@pytest.mark.parametrize("uniform", [False, True])
@pytest.mark.parametrize("strategy", [dk.FullPrecision, dk.Hybrid])
def test_large_mixed_batch_equals_small_launches(strategy, uniform):
    # uniform: every query has the same L (the launch keeps its own order)
    log("ordered launch strategy=%d uniform=%d" % (strategy, uniform))
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    piv, off = pq_setup(vec, 32)
    ix.set_pq(piv, off)
    ix.encode()
    n = 3000
    rng = np.random.default_rng(9)
    q = (qs[rng.integers(0, qs.shape[0], n)] +
         0.3 * rng.standard_normal((n, vec.shape[1]))).astype(np.float32)
    ks, Ls = mixed(n, 4)
    if uniform:
        Ls[:] = 64
    big = ix.search_batch(q, ks, Ls, 2, strategy)
    ids = np.zeros_like(big.ids)
    ds = np.zeros_like(big.distances)
    cnt = np.zeros_like(big.counts)
    cm = np.zeros_like(big.cmps)
    hp = np.zeros_like(big.hops)
    for a in range(0, n, 500):
        r = ix.search_batch(q[a:a + 500], ks[a:a + 500], Ls[a:a + 500], 2, strategy)
        w = r.ids.shape[1]
        ids[a:a + 500, :w], ds[a:a + 500, :w] = r.ids, r.distances
        cnt[a:a + 500], cm[a:a + 500], hp[a:a + 500] = r.counts, r.cmps, r.hops
    check_same(big, dk.SearchResult(ids, ds, cnt, cm, hp), n)
    with dk.SearchGraph(ix, q, ks, Ls, 2, strategy) as g:
        for _ in range(2):
            check_same(g.run(), big, n)
            ix.search_batch(q[:64], ks[:64], Ls[:64], 2, strategy)  # eager work in between
