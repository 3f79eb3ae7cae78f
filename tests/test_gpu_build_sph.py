"""GPU 1-bit spherical build (dkm_set_build_spherical: Quantized::build() over a
SphericalStore, diskann-providers/src/model/graph/provider/async_/inmem/spherical.rs:293-350,
515-533, 605-650) against the oracle's spherical build on the same data and quantizer:
identical adjacency lists, bit for bit. The insert search traverses with SameAsData queries
and no inner-product rescale; every prune and back-edge distance is the data x data distance
of two stored codes. Also dkm_prune_batch under the strategy and its preconditions."""
import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd import spherical
from diskann_amd.errors import ANNError
from oracle import pyoracle as po
from test_gpu_search import log

pytestmark = pytest.mark.gpu


def data(metric, n, dim, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((8, dim)).astype(np.float32) * 3
    x = (c[rng.integers(0, 8, n)] + rng.standard_normal((n, dim))).astype(np.float32)
    if metric == po.COSINE:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    med, _ = po.medoid(x)
    return np.vstack([x, x[med:med + 1]])


def sph_index(metric, vec, R, R_max, layout="four_bit_transposed"):
    n = vec.shape[0] - 1
    g = spherical.SphericalQuantizer.train(vec[:-1], metric, seed=vec.shape[1])
    ix = dk.DiskANNIndex(metric, vec.shape[1], n, R, R_max)
    ix.set_vectors(vec[:n])
    ix.set_start_points(vec[n:])
    ix.set_spherical(g, layout)
    ix.encode_spherical()
    o = po.Spherical(metric, g.centroid, g.mean_norm, 1.0, po.SPH_DOUBLE_HADAMARD, g.signs0,
                     g.signs1)
    stored = ix.get_sph_codes()
    assert np.array_equal(stored, o.compress(vec))
    return ix, o, stored


@pytest.mark.parametrize("metric,dim", [(po.L2, 64), (po.INNER_PRODUCT, 48), (po.COSINE, 96)])
@pytest.mark.parametrize("batch,intra", [(1, 0), (16, -1)])
def test_sph_build_matches_oracle(metric, dim, batch, intra):
    log("sph build metric=%d dim=%d batch=%d intra=%d" % (metric, dim, batch, intra))
    n, R, R_max, L = 700, 16, 20, 32
    vec = data(metric, n, dim, seed=dim)
    ix, o, stored = sph_index(metric, vec, R, R_max)
    ix.set_build_spherical(True)
    ix.build(np.arange(n), batch=batch, intra_batch=intra, L_build=L, alpha=1.2)
    adj_o, deg_o = po.build_graph(metric, vec, R=R, R_max=R_max, L_build=L, alpha=1.2,
                                  starts=[n], ids=np.arange(n), batch=batch, intra_batch=intra,
                                  sph=(o, stored))
    lists, counts = ix.get_graph()
    assert np.array_equal(counts, deg_o)
    for i in range(n + 1):
        assert list(lists[i, :counts[i]]) == list(adj_o[i, :deg_o[i]]), i


def test_sph_prune_batch_matches_oracle():
    log("sph prune batch")
    n, dim, R = 500, 64, 12
    vec = data(po.L2, n, dim, seed=3)
    ix, o, stored = sph_index(po.L2, vec, R, 16)
    ix.set_build_spherical(True)
    rng = np.random.default_rng(5)
    locs = np.arange(0, n, 25, dtype=np.uint32)
    lists = [rng.choice(n, 90, replace=False).astype(np.uint32) for _ in locs]
    got = ix.prune_batch(locs, lists, alpha=1.2)
    for j, loc in enumerate(locs):
        want = po.prune_list(po.L2, vec, int(loc), lists[j], R, 1.2, sph=(o, stored))
        assert list(got[j]) == list(want), j


def test_sph_build_preconditions():
    log("sph build preconditions")
    vec = data(po.L2, 100, 16)
    ix = dk.DiskANNIndex(po.L2, 16, 100, 8, 10)
    ix.set_vectors(vec[:100])
    ix.set_start_points(vec[100:])
    ix.set_build_spherical(True)
    with pytest.raises(ANNError):           # no spherical quantizer yet
        ix.build(np.arange(100), batch=1, L_build=16)
    ix.set_build_sq(True)
    with pytest.raises(ANNError):           # the two quantized builds exclude each other
        ix.build(np.arange(100), batch=1, L_build=16)
