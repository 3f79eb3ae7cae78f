"""GPU scalar-quantized build (dkm_set_build_sq: the Quantized strategy over an SQStore,
diskann-providers/src/model/graph/provider/async_/inmem/scalar.rs:370-446, 595-720) against
the oracle's SQ build on the same data and quantizer: identical adjacency lists, bit for bit.
The insert search traverses with the SQ query computer (inner-product queries rescaled to
the mean norm) and every prune and back-edge distance is the compensated distance between
stored codes. Single inserts are what the reference's SQ benchmark runs; batches use the
same distances. Also: dkm_prune_batch under the SQ strategy == the oracle's
robust_prune_list with SQ distances, and the strategy's preconditions."""
import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd import scalar
from diskann_amd.errors import ANNError
from oracle import pyoracle as po
from test_gpu_search import log

pytestmark = pytest.mark.gpu


def data(metric, n, dim, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((8, dim)).astype(np.float32) * 3
    x = (c[rng.integers(0, 8, n)] + rng.standard_normal((n, dim))).astype(np.float32)
    if metric != po.L2:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    med, _ = po.medoid(x)
    return np.vstack([x, x[med:med + 1]])


def sq_build_index(metric, vec, R, R_max, L, batch, intra):
    n = vec.shape[0] - 1
    qz = scalar.ScalarQuantizer.train(vec[:-1])
    ix = dk.DiskANNIndex(metric, vec.shape[1], n, R, R_max)
    ix.set_vectors(vec[:n])
    ix.set_start_points(vec[n:])
    ix.set_sq(qz)
    ix.encode_sq()
    ix.set_build_sq(True)
    ix.build(np.arange(n), batch=batch, intra_batch=intra, L_build=L, alpha=1.2)
    return ix, qz


@pytest.mark.parametrize("metric,dim", [(po.L2, 32), (po.L2, 96), (po.INNER_PRODUCT, 48),
                                        (po.COSINE_NORMALIZED, 64)])
@pytest.mark.parametrize("batch,intra", [(1, 0), (16, -1), (64, 0)])
def test_sq_build_matches_oracle(metric, dim, batch, intra):
    log("sq build metric=%d dim=%d batch=%d intra=%d" % (metric, dim, batch, intra))
    n, R, R_max, L = 700, 16, 20, 32
    vec = data(metric, n, dim, seed=dim)
    ix, qz = sq_build_index(metric, vec, R, R_max, L, batch, intra)
    oc, ocomp = po.sq_compress(qz.scale, qz.shift, vec)
    codes, comp = scalar.split(ix.get_sq_codes())
    assert np.array_equal(codes, oc) and comp.tobytes() == ocomp.tobytes()
    adj_o, deg_o = po.build_graph(metric, vec, R=R, R_max=R_max, L_build=L, alpha=1.2,
                                  starts=[n], ids=np.arange(n), batch=batch, intra_batch=intra,
                                  sq=(qz.scale, qz.shift, qz.mean_norm, oc, ocomp))
    lists, counts = ix.get_graph()
    assert np.array_equal(counts, deg_o)
    for i in range(n + 1):
        assert list(lists[i, :counts[i]]) == list(adj_o[i, :deg_o[i]]), i
    # the SQ graph differs from the full-precision one: the strategy changed the distances
    adj_f, deg_f = po.build_graph(metric, vec, R=R, R_max=R_max, L_build=L, alpha=1.2,
                                  starts=[n], ids=np.arange(n), batch=batch, intra_batch=intra)
    assert any(list(adj_f[i, :deg_f[i]]) != list(adj_o[i, :deg_o[i]]) for i in range(n))


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT])
def test_sq_prune_batch_matches_oracle(metric):
    log("sq prune batch metric=%d" % metric)
    n, dim, R = 500, 40, 12
    vec = data(metric, n, dim, seed=3)
    qz = scalar.ScalarQuantizer.train(vec[:-1])
    ix = dk.DiskANNIndex(metric, dim, n, R, 16)
    ix.set_vectors(vec[:n])
    ix.set_start_points(vec[n:])
    ix.set_sq(qz)
    ix.encode_sq()
    ix.set_build_sq(True)
    rng = np.random.default_rng(5)
    locs = np.arange(0, n, 25, dtype=np.uint32)
    lists = np.stack([rng.choice(n, 90, replace=False) for _ in locs]).astype(np.uint32)
    got = ix.prune_batch(locs, list(lists), alpha=1.2)
    oc, ocomp = po.sq_compress(qz.scale, qz.shift, vec)
    sq = (qz.scale, qz.shift, qz.mean_norm, oc, ocomp)
    for j, loc in enumerate(locs):
        want = po.prune_list(metric, vec, int(loc), lists[j], R, 1.2, sq=sq)
        assert list(got[j]) == list(want), j


def test_sq_build_preconditions():
    log("sq build preconditions")
    vec = data(po.L2, 100, 16)
    ix = dk.DiskANNIndex(po.L2, 16, 100, 8, 10)
    ix.set_vectors(vec[:100])
    ix.set_start_points(vec[100:])
    ix.set_build_sq(True)
    with pytest.raises(ANNError):           # no scalar quantizer yet
        ix.build(np.arange(100), batch=1, L_build=16)
    ix.set_sq(scalar.ScalarQuantizer.train(vec[:100]))
    with pytest.raises(ANNError):           # codes not encoded
        ix.build(np.arange(100), batch=1, L_build=16)
