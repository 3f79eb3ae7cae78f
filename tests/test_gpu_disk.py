"""GPU disk-index search (dkm_disk_search_batch) against the oracle's DiskAccessor +
RerankAndFilter restatement (diskann-disk/src/search/provider/disk_provider.rs:323-388,
666-721): ids, f32 distance bytes, cmps and hops, for the transposed PQ table with the
lookup table in LDS and with direct distances, every metric, beam widths 1 / 2 / 4,
mixed k / L and I/O budgets."""
import numpy as np
import pytest

import diskann_amd as dk
from oracle import pyoracle as po
from test_gpu_search import dataset, log, pq_setup

pytestmark = pytest.mark.gpu


def disk_index(metric, dim, M):
    vec, adj, deg, qs = dataset(metric, dim=dim)
    n = vec.shape[0] - 1
    x, adj, deg = vec[:n], adj[:n], deg[:n].copy()
    med, _ = po.medoid(x)
    # the disk graph holds the data rows only: the frozen row's edges are dropped
    adj = np.where(adj >= n, med, adj).astype(np.uint32)
    for i in range(n):
        row = list(dict.fromkeys(int(v) for v in adj[i, :deg[i]] if v != i))
        adj[i, :len(row)] = row
        deg[i] = len(row)
    piv, off = pq_setup(vec, M)
    ix = dk.DiskANNIndex(metric, dim, n, adj.shape[1], adj.shape[1])
    ix.set_vectors(x)
    ix.set_start_points(x[med])
    ix.set_graph(adj, deg)
    ix.set_pq(piv, off)
    ix.encode()
    codes = ix.get_codes()[:n]
    oix = po.OracleIndex(metric, x, adj, deg, [med], codes, piv, off)
    return ix, oix, med, qs


@pytest.mark.parametrize("metric,dim,M", [(po.L2, 96, 48), (po.L2, 128, 32), (po.L2, 40, 7),
                                          (po.INNER_PRODUCT, 64, 16), (po.COSINE, 48, 12)])
@pytest.mark.parametrize("adc_mode", [0, 1])
def test_disk_search_matches_oracle(metric, dim, M, adc_mode):
    log("disk search metric=%d dim=%d M=%d adc=%d" % (metric, dim, M, adc_mode))
    ix, oix, med, qs = disk_index(metric, dim, M)
    ix.set_tuning(0, adc_mode)
    rng = np.random.default_rng(dim + M)
    ks = rng.choice([1, 5, 10], qs.shape[0]).astype(np.uint32)
    Ls = np.maximum(ks, rng.integers(8, 96, qs.shape[0])).astype(np.uint32)
    for bw in (1, 2, 4):
        r = ix.disk_search_batch(qs, ks, Ls, bw, med)
        for i in range(qs.shape[0]):
            ids, ds, cm, hp = oix.search_disk(qs[i], int(ks[i]), int(Ls[i]), bw)
            assert r.counts[i] == len(ids), (bw, i)
            assert list(r.ids[i, :len(ids)]) == list(ids), (bw, i)
            assert r.distances[i, :len(ids)].tobytes() == ds.tobytes(), (bw, i)
            assert (r.cmps[i], r.hops[i]) == (cm, hp), (bw, i)


@pytest.mark.parametrize("io_limit", [0, 1, 5, 30])
def test_disk_search_io_budget_matches_oracle(io_limit):
    log("disk search io_limit=%d" % io_limit)
    ix, oix, med, qs = disk_index(po.L2, 96, 48)
    for bw in (1, 4):
        r = ix.disk_search_batch(qs, 10, 64, bw, med, io_limit=io_limit)
        for i in range(qs.shape[0]):
            ids, ds, cm, hp = oix.search_disk(qs[i], 10, 64, bw, io_limit=io_limit)
            assert list(r.ids[i, :r.counts[i]]) == list(ids), (bw, i)
            assert r.distances[i, :len(ids)].tobytes() == ds.tobytes(), (bw, i)
            assert (r.cmps[i], r.hops[i]) == (cm, hp), (bw, i)


def test_disk_search_cosine_normalized_matches_oracle():
    vec, adj, deg, qs = dataset(po.INNER_PRODUCT, dim=64)     # unit-norm rows
    n = vec.shape[0] - 1
    x = vec[:n]
    med, _ = po.medoid(x)
    adj = np.where(adj[:n] >= n, med, adj[:n]).astype(np.uint32)
    deg = deg[:n].copy()
    for i in range(n):
        row = list(dict.fromkeys(int(v) for v in adj[i, :deg[i]] if v != i))
        adj[i, :len(row)] = row
        deg[i] = len(row)
    piv, off = pq_setup(vec, 16)
    ix = dk.DiskANNIndex(po.COSINE_NORMALIZED, 64, n, adj.shape[1], adj.shape[1])
    ix.set_vectors(x)
    ix.set_start_points(x[med])
    ix.set_graph(adj, deg)
    ix.set_pq(piv, off)
    ix.encode()
    oix = po.OracleIndex(po.COSINE_NORMALIZED, x, adj, deg, [med], ix.get_codes()[:n], piv, off)
    r = ix.disk_search_batch(qs, 10, 50, 2, med)
    for i in range(qs.shape[0]):
        ids, ds, cm, hp = oix.search_disk(qs[i], 10, 50, 2)
        assert list(r.ids[i, :r.counts[i]]) == list(ids), i
        assert r.distances[i, :len(ids)].tobytes() == ds.tobytes(), i
        assert (r.cmps[i], r.hops[i]) == (cm, hp), i


def test_disk_search_errors():
    ix, oix, med, qs = disk_index(po.L2, 96, 48)
    with pytest.raises(dk.KnnSearchError):
        ix.disk_search_batch(qs[:2], 10, 5, 1, med)        # L < k
    with pytest.raises(dk.ANNError):
        ix.disk_search_batch(qs[:2], 5, 10, 1, 10 ** 6)    # medoid not a data row
