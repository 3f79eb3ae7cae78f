"""Standalone RobustPrune on the GPU (dkm_prune_batch, dkm_prune_range) against the
oracle's robust_prune_list (diskann/src/graph/index.rs:2402-2450, occlude_list
:2570-2782) and prune_range (:2788-2830): the pruned lists must be identical id for id,
for every metric, alpha, saturation, pool truncation and the Hybrid prune."""
import numpy as np
import pytest

from oracle import pyoracle as po
from test_gpu_search import dataset, log, make_index, pq_setup

pytestmark = pytest.mark.gpu


def candidate_lists(vec, adj, deg, nloc, rng, width):
    """Per location: its graph neighbours, their neighbours and random rows (distinct,
    the location itself kept in to be dropped by the prune)."""
    n = vec.shape[0]
    locs = rng.choice(n, nloc, replace=False).astype(np.uint32)
    lists = []
    for v in locs:
        c = list(adj[v, :deg[v]])
        for u in adj[v, :deg[v]][:8]:
            c += list(adj[u, :deg[u]])
        c += list(rng.integers(0, n, width // 4))
        c.append(int(v))
        seen, out = set(), []
        for x in c:
            if int(x) not in seen:
                seen.add(int(x))
                out.append(int(x))
        rng.shuffle(out)
        lists.append(np.array(out[:width], np.uint32))
    return locs, lists


@pytest.mark.parametrize("metric,dim", [(po.L2, 128), (po.INNER_PRODUCT, 64), (po.COSINE, 48),
                                        (po.COSINE_NORMALIZED, 96)])
@pytest.mark.parametrize("alpha,saturate,force", [(1.2, False, False), (1.0, False, False),
                                                  (1.2, True, False), (1.0, False, True)])
def test_prune_batch_matches_oracle(metric, dim, alpha, saturate, force):
    log("prune batch metric=%d dim=%d alpha=%.1f sat=%d force=%d" % (metric, dim, alpha,
                                                                     saturate, force))
    if metric == po.COSINE_NORMALIZED:
        vec, adj, deg, _ = dataset(po.INNER_PRODUCT, dim=dim)   # unit-norm rows
    else:
        vec, adj, deg, _ = dataset(metric, dim=dim)
    ix = make_index(metric, vec, adj, deg)
    rng = np.random.default_rng(dim)
    locs, lists = candidate_lists(vec, adj, deg, 200, rng, 300)
    got = ix.prune_batch(locs, lists, alpha=alpha, saturate=saturate, force_saturate=force)
    for i, v in enumerate(locs):
        want = po.prune_list(metric, vec, v, lists[i], ix.R, alpha, saturate=saturate,
                             force_saturate=force)
        assert np.array_equal(got[i], want), i
        assert int(v) not in set(got[i].tolist())


def test_prune_batch_truncated_pool_and_edge_lists():
    vec, adj, deg, _ = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    rng = np.random.default_rng(7)
    locs, lists = candidate_lists(vec, adj, deg, 64, rng, 400)
    # short lists, a list of only the location itself and an empty list
    lists[0] = lists[0][:3]
    lists[1] = np.array([locs[1]], np.uint32)
    lists[2] = np.array([], np.uint32)
    got = ix.prune_batch(locs, lists, alpha=1.2, max_occlusion=60)
    for i, v in enumerate(locs):
        want = po.prune_list(po.L2, vec, v, lists[i], ix.R, 1.2, max_occlusion=60)
        assert np.array_equal(got[i], want), i
    assert len(got[1]) == 0 and len(got[2]) == 0


@pytest.mark.parametrize("max_fp", [0, 16, None])
def test_prune_batch_hybrid_matches_oracle(max_fp):
    log("prune batch hybrid max_fp=%s" % max_fp)
    vec, adj, deg, _ = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    piv, off = pq_setup(vec, 32)
    ix.set_pq(piv, off)
    ix.encode()
    codes = ix.get_codes()
    ix.set_build_hybrid(True, max_fp)
    rng = np.random.default_rng(11)
    locs, lists = candidate_lists(vec, adj, deg, 100, rng, 250)
    got = ix.prune_batch(locs, lists, alpha=1.2)
    mf = 1 << 30 if max_fp is None else max_fp
    for i, v in enumerate(locs):
        want = po.prune_list(po.L2, vec, v, lists[i], ix.R, 1.2, max_fp=mf, codes=codes,
                             pivots=piv, pq_offsets=off)
        assert np.array_equal(got[i], want), i


def test_prune_range_matches_oracle():
    log("prune range")
    vec, adj, deg, _ = dataset(po.L2)
    R = 16            # below the built lists' degrees (up to 31): most lists get pruned
    ix = make_index(po.L2, vec, adj, deg, R=R)
    n = vec.shape[0] - 1
    ids = np.arange(0, n, 3, dtype=np.uint32)
    ix.prune_range(ids, alpha=1.2)
    got_adj, got_deg = ix.get_graph()
    touched = set(ids.tolist())
    for v in range(n):
        if v in touched and deg[v] > R:
            want = po.prune_list(po.L2, vec, v, adj[v, :deg[v]], R, 1.2)
        else:
            want = adj[v, :deg[v]]
        assert np.array_equal(got_adj[v, :got_deg[v]], want), v


def test_prune_errors():
    vec, adj, deg, _ = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    import diskann_amd as dk
    with pytest.raises(dk.ANNError):
        ix.prune_batch([0], [np.array([vec.shape[0] + 5], np.uint32)])    # id out of range
    with pytest.raises(dk.ANNError):
        ix.prune_batch([0], [np.array([1, 2], np.uint32)], alpha=0.5)
    with pytest.raises(dk.ANNError):
        ix.prune_range([vec.shape[0] + 5])
