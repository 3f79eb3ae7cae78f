"""GPU 8-bit scalar quantization (training, compression, compensated distances, and the
scalar-quantized traversal with and without the FP32 rerank) against the oracle on the
same inputs: every integer output and every f32 value bit-for-bit."""
import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd import scalar
from oracle import pyoracle as po
from test_gpu_search import dataset, log, make_index

pytestmark = pytest.mark.gpu


def rows(n, dim, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, dim)) * rng.uniform(0.2, 3.0, dim)
            + rng.uniform(-4, 4, dim)).astype(np.float32)


@pytest.mark.parametrize("n,dim", [(1, 5), (7, 1), (1000, 128), (333, 300), (2000, 37)])
@pytest.mark.parametrize("stdevs", [1.0, 2.0])
def test_train_matches_oracle(n, dim, stdevs):
    log("sq train n=%d dim=%d stdevs=%g" % (n, dim, stdevs))
    x = rows(n, dim, n + dim)
    qz = scalar.ScalarQuantizer.train(x, stdevs)
    scale, shift, mean_norm = po.sq_train(x, stdevs)
    assert np.float32(qz.scale) == np.float32(scale)
    assert qz.shift.tobytes() == shift.tobytes()
    assert np.float32(qz.mean_norm) == np.float32(mean_norm)


@pytest.mark.parametrize("dim", [1, 5, 16, 37, 128, 300, 1536])
def test_compress_matches_oracle(dim):
    log("sq compress dim=%d" % dim)
    x = rows(150, dim, dim)
    qz = scalar.ScalarQuantizer.train(x[:100])
    out = qz.compress(x)          # rows 100..149 reach outside the trained range too
    codes, comp = scalar.split(out)
    oc, ocomp = po.sq_compress(qz.scale, qz.shift, x)
    assert np.array_equal(codes, oc)
    assert comp.tobytes() == ocomp.tobytes()


def test_compress_nan_is_an_error():
    x = rows(3, 8, 1)
    qz = scalar.ScalarQuantizer.train(x)
    x[1, 3] = np.nan
    with pytest.raises(dk.ANNError):
        qz.compress(x)


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE_NORMALIZED])
@pytest.mark.parametrize("dim", [3, 16, 128, 333, 1536])
def test_distances_match_oracle(metric, dim):
    log("sq distance metric=%d dim=%d" % (metric, dim))
    x = rows(60, dim, 2 * dim + metric)
    qz = scalar.ScalarQuantizer.train(x)
    cx = qz.compress(x)
    rng = np.random.default_rng(dim)
    ids = rng.integers(0, 60, (12, 17)).astype(np.uint32)
    got = qz.distance(metric, cx[:12], cx, ids)
    codes, comp = scalar.split(cx)
    ssn = po.sq_shift_square_norm(qz.shift)
    for i in range(12):
        for j in range(17):
            r = ids[i, j]
            e = po.sq_distance(metric, qz.scale, ssn, codes[i], comp[i], codes[r], comp[r])
            assert got[i, j] == np.float32(e), (i, j)


def test_distance_errors():
    x = rows(10, 8, 3)
    qz = scalar.ScalarQuantizer.train(x)
    cx = qz.compress(x)
    with pytest.raises(dk.ANNError):
        qz.distance(po.COSINE, cx[:1], cx, np.zeros((1, 1), np.uint32))
    with pytest.raises(dk.ANNError):
        qz.distance(po.L2, cx[:1], cx, np.array([[10]], np.uint32))


def sq_index(metric, dim, n=3000, R=24):
    vec, adj, deg, qs = dataset(metric, n=n, dim=dim, R=R)
    qz = scalar.ScalarQuantizer.train(vec[:-1])
    ix = make_index(metric, vec, adj, deg)
    ix.set_sq(qz)
    ix.encode_sq()
    return ix, qz, vec, adj, deg, qs


@pytest.mark.parametrize("metric,dim,n", [(po.L2, 128, 3000), (po.INNER_PRODUCT, 64, 3000),
                                          (po.COSINE_NORMALIZED, 96, 3000),
                                          (po.L2, 300, 1500), (po.INNER_PRODUCT, 600, 1200)])
@pytest.mark.parametrize("strategy", [dk.Hybrid, dk.Quantized])
@pytest.mark.parametrize("L,bw", [(32, 1), (64, 2)])
def test_sq_search_matches_oracle(metric, dim, n, strategy, L, bw):
    log("sq search metric=%d dim=%d strat=%d L=%d bw=%d" % (metric, dim, strategy, L, bw))
    ix, qz, vec, adj, deg, qs = sq_index(metric, dim, n=n)
    stored = ix.get_sq_codes()
    codes, comp = scalar.split(stored)
    oc, ocomp = po.sq_compress(qz.scale, qz.shift, vec)
    assert np.array_equal(codes, oc) and comp.tobytes() == ocomp.tobytes()
    oix = po.OracleIndex(metric, vec, adj, deg, [vec.shape[0] - 1],
                         sq=(qz.scale, qz.shift, qz.mean_norm, oc, ocomp))
    mode = po.MODE_HYBRID if strategy == dk.Hybrid else po.MODE_PQ_ONLY
    r = ix.search_batch(qs, 10, L, bw, strategy)
    for i in range(qs.shape[0]):
        ids, ds, cm, hp = oix.search(qs[i], 10, L, bw, mode)
        assert list(r.ids[i, :len(ids)]) == list(ids), i
        assert r.distances[i, :len(ids)].tobytes() == ds.tobytes(), i
        assert (r.cmps[i], r.hops[i]) == (cm, hp), i


def test_sq_codes_round_trip_and_switch_back_to_pq():
    ix, qz, vec, adj, deg, qs = sq_index(po.L2, 128)
    stored = ix.get_sq_codes()
    ix2 = make_index(po.L2, vec, adj, deg)
    ix2.set_sq(qz)
    with pytest.raises(dk.ANNError):
        ix2.search_batch(qs[:2], 10, 32, 1, dk.Hybrid)     # codes not set yet
    ix2.set_sq_codes(stored)
    a = ix.search_batch(qs, 10, 40, 1, dk.Hybrid)
    b = ix2.search_batch(qs, 10, 40, 1, dk.Hybrid)
    assert np.array_equal(a.ids, b.ids) and a.distances.tobytes() == b.distances.tobytes()
    ix2.use_pq()
    with pytest.raises(dk.ANNError):
        ix2.search_batch(qs[:2], 10, 32, 1, dk.Hybrid)     # no PQ store


def test_sq_rejects_cosine_index():
    vec, adj, deg, qs = dataset(po.COSINE, dim=48)
    ix = make_index(po.COSINE, vec, adj, deg)
    with pytest.raises(dk.ANNError):
        ix.set_sq(scalar.ScalarQuantizer.train(vec[:-1]))
