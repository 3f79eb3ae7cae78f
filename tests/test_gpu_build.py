"""GPU graph construction (multi_insert batches: search + RobustPrune + back edges)
against the oracle's build on the same inputs: identical adjacency lists, bit for bit;
plus the reference's golden grid_insert cases through the GPU build."""
import json
import os

import numpy as np
import pytest

import diskann_amd as dk
from gridutil import grid
from oracle import pyoracle as po
from test_oracle_build import case_params

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "grid_cases.json")))


def gpu_build(metric, vec, R, R_max, L_build, alpha, batch, intra, nstart=1):
    n = vec.shape[0] - nstart
    ix = dk.DiskANNIndex(metric, vec.shape[1], n, R, R_max, num_start=nstart)
    ix.set_vectors(vec[:n])
    ix.set_start_points(vec[n:])
    ix.build(np.arange(n), batch=batch, intra_batch=intra, L_build=L_build, alpha=alpha)
    return ix


def data(metric, n, dim, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((8, dim)).astype(np.float32) * 3
    x = (c[rng.integers(0, 8, n)] + rng.standard_normal((n, dim))).astype(np.float32)
    if metric != po.L2:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    med, _ = po.medoid(x)
    return np.vstack([x, x[med:med + 1]])


@pytest.mark.parametrize("metric,dim", [(po.L2, 32), (po.L2, 128), (po.INNER_PRODUCT, 24),
                                        (po.COSINE, 20)])
@pytest.mark.parametrize("batch,intra", [(1, 0), (16, -1), (64, 0), (64, 8), (200, 0)])
def test_gpu_build_matches_oracle(metric, dim, batch, intra):
    print("build metric=%d dim=%d batch=%d intra=%d" % (metric, dim, batch, intra), flush=True)
    n, R, R_max, L = 800, 16, 20, 32
    vec = data(metric, n, dim)
    adj_o, deg_o = po.build_graph(metric, vec, R=R, R_max=R_max, L_build=L, alpha=1.2,
                                  starts=[n], ids=np.arange(n), batch=batch, intra_batch=intra)
    ix = gpu_build(metric, vec, R, R_max, L, 1.2, batch, intra)
    lists, counts = ix.get_graph()
    assert np.array_equal(counts, deg_o)
    for i in range(n + 1):
        assert list(lists[i, :counts[i]]) == list(adj_o[i, :deg_o[i]]), i


@pytest.mark.parametrize("name", sorted(FIX["grid_insert"]))
def test_gpu_build_grid_insert_golden(name):
    c = FIX["grid_insert"][name]
    dim, size, batch, intra = case_params(name)
    vec, _, _, starts = grid(dim, size)
    R_max = 2 * dim
    R = min(max(2, R_max - 2), R_max)
    ix = gpu_build(po.L2, vec, R, R_max, 100, 1.2, batch, intra)
    for s in c["searches"]:
        ids, ds, cm, hp = ix.search(dk.Knn(10, 10), dk.FullPrecision,
                                    np.array(s["query"], np.float32))
        assert [float(d) for d in ds] == [float(e[1]) for e in s["results"]], name
        if dim == 1:
            assert [int(i) for i in ids] == [int(e[0]) for e in s["results"]], name
            assert (cm, hp) == (s["comparisons"], s["hops"]), name
