"""GPU 1-bit spherical quantization against the reference's golden baselines and the
oracle: compressed bytes, data and query distances of every layout, training
statistics, and the 1-bit traversal (same-as-data and 4-bit query layouts, with and
without the FP32 rerank) - all bit for bit."""
import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd import spherical
import sphutil
from oracle import pyoracle as po
from test_gpu_search import dataset, log, make_index

pytestmark = pytest.mark.gpu
G = sphutil.load()
X = np.array(G["dataset"], np.float32)
KIND = {"null": po.SPH_NULL, "padding_hadamard": po.SPH_PADDING_HADAMARD,
        "double_hadamard": po.SPH_DOUBLE_HADAMARD}


@pytest.mark.parametrize("metric", sorted(sphutil.METRICS))
def test_golden_baseline(metric):
    log("sph golden %s" % metric)
    b = G["baselines"][metric]
    q = spherical.SphericalQuantizer.from_serialized(b["serialized_quantizer"])
    codes = q.compress(X)
    assert np.array_equal(codes, np.array(b["compressed_vectors"], np.uint8))
    ii, jj = np.triu_indices(16)
    ids = np.zeros((16, 16), np.uint32)
    ids[:] = np.arange(16)
    dd = q.distance(codes, codes, ids)
    assert np.array_equal(dd[ii, jj], np.array(b["data_distances"], np.float32))
    keys = ["query_distances"] + (["rescaled_query_distances"]
                                  if "rescaled_query_distances" in b else [])
    for key in keys:
        for L in b[key]:
            got = q.query_distance(X, codes, ids, L["layout"], key.startswith("rescaled"))
            assert np.array_equal(got, np.array(L["distances"], np.float32)), (key, L["layout"])


@pytest.mark.parametrize("metric", sorted(sphutil.METRICS))
def test_train_stats_match_golden(metric):
    b = G["baselines"][metric]
    f = spherical.parse_serialized(b["serialized_quantizer"])
    q = spherical.SphericalQuantizer.train(X, metric)
    assert np.array_equal(q.centroid, f["centroid"])


def oracle_and_gpu(metric, dim, kind, sub, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((64, dim)) + rng.uniform(-1, 1, dim)).astype(np.float32)
    c, mn = po.sph_train_stats(metric, x)
    s0 = rng.integers(0, 2, dim).astype(bool)
    padded = 1 << (dim - 1).bit_length()
    if kind == "double_hadamard":
        out = dim if sub is None else sub
        n = max(dim, out)
        s1 = rng.integers(0, 2, n).astype(bool)
        subs = None if sub is None else np.sort(rng.choice(dim, sub, replace=False))
        o = po.Spherical(metric, c, mn, 1.0, po.SPH_DOUBLE_HADAMARD, s0, s1, subsample=subs)
        g = spherical.SphericalQuantizer(metric, c, mn, 1.0, kind, s0, s1, 0, subs)
    elif kind == "padding_hadamard":
        subs = None if sub is None else np.sort(rng.choice(padded, sub, replace=False))
        o = po.Spherical(metric, c, mn, 1.0, po.SPH_PADDING_HADAMARD, s0, None, padded, subs)
        g = spherical.SphericalQuantizer(metric, c, mn, 1.0, kind, s0, None, padded, subs)
    else:
        o = po.Spherical(metric, c, mn, 1.0, po.SPH_NULL)
        g = spherical.SphericalQuantizer(metric, c, mn, 1.0, "null")
    return x, o, g


CASES = [("double_hadamard", 37, None), ("double_hadamard", 128, None),
         ("double_hadamard", 300, None), ("double_hadamard", 200, 150),
         ("padding_hadamard", 100, None), ("padding_hadamard", 768, None),
         ("padding_hadamard", 100, 90), ("null", 24, None), ("double_hadamard", 1536, None)]


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE])
@pytest.mark.parametrize("kind,dim,sub", CASES)
def test_compress_and_distances_match_oracle(metric, kind, dim, sub):
    log("sph oracle metric=%d %s dim=%d sub=%s" % (metric, kind, dim, sub))
    x, o, g = oracle_and_gpu(metric, dim, kind, sub, dim + metric)
    assert g.out_dim == o.out_dim
    codes = g.compress(x)
    assert np.array_equal(codes, o.compress(x))
    rng = np.random.default_rng(dim)
    ids = rng.integers(0, 64, (8, 12)).astype(np.uint32)
    dd = g.distance(codes[:8], codes, ids)
    for i in range(8):
        for j in range(12):
            assert dd[i, j] == np.float32(o.dist_data(codes[i], codes[ids[i, j]])), (i, j)
    qs = x[:8] * np.float32(0.7) + np.float32(0.1)
    for layout in ("same_as_data", "four_bit_transposed", "full_precision"):
        for resc in (False, True):
            got = g.query_distance(qs, codes, ids, layout, resc)
            for i in range(8):
                qv = o.rescale(qs[i]) if (resc and metric == po.INNER_PRODUCT) else qs[i]
                if layout == "same_as_data":
                    qc = o.compress(qv)[0]
                    exp = [o.dist_data(qc, codes[r]) for r in ids[i]]
                elif layout == "four_bit_transposed":
                    c4, m4 = o.compress_query4(qv)
                    exp = [o.dist_query4(c4, m4, codes[r]) for r in ids[i]]
                else:
                    v, m = o.compress_full(qv)
                    exp = [o.dist_full(v, m, codes[r]) for r in ids[i]]
                assert np.array_equal(got[i], np.array(exp, np.float32)), (layout, resc, i)


def test_centroid_row_gets_default_meta():
    x, o, g = oracle_and_gpu(po.L2, 64, "double_hadamard", None, 5)
    c = g.compress(g.centroid[None, :])
    assert np.array_equal(c, o.compress(o.shift[None, :]))
    assert not c.any()


def test_errors():
    x, o, g = oracle_and_gpu(po.L2, 32, "double_hadamard", None, 6)
    bad = x[:2].copy()
    bad[1, 3] = np.nan
    with pytest.raises(dk.ANNError):
        g.compress(bad)
    with pytest.raises(dk.ANNError):
        spherical.SphericalQuantizer(po.L2, np.zeros(8, np.float32), 1.0, 1.0,
                                     "padding_hadamard", np.zeros(8), None, 12)


def sph_index(metric, dim, layout, n=3000):
    vec, adj, deg, qs = dataset(metric, n=n, dim=dim)
    g = spherical.SphericalQuantizer.train(vec[:-1], metric, seed=dim)
    ix = make_index(metric, vec, adj, deg)
    ix.set_spherical(g, layout)
    ix.encode_spherical()
    return ix, g, vec, adj, deg, qs


@pytest.mark.parametrize("metric,dim", [(po.L2, 128), (po.INNER_PRODUCT, 64),
                                        (po.COSINE, 48), (po.L2, 300)])
@pytest.mark.parametrize("layout", ["four_bit_transposed", "same_as_data"])
@pytest.mark.parametrize("strategy", [dk.Hybrid, dk.Quantized])
@pytest.mark.parametrize("L,bw", [(32, 1), (64, 2)])
def test_sph_search_matches_oracle(metric, dim, layout, strategy, L, bw):
    log("sph search metric=%d dim=%d %s strat=%d L=%d bw=%d" % (metric, dim, layout, strategy,
                                                                 L, bw))
    ix, g, vec, adj, deg, qs = sph_index(metric, dim, layout, n=3000 if dim <= 128 else 1500)
    stored = ix.get_sph_codes()
    o = po.Spherical(metric, g.centroid, g.mean_norm, 1.0, po.SPH_DOUBLE_HADAMARD,
                     g.signs0, g.signs1)
    assert np.array_equal(stored, o.compress(vec))
    oix = po.OracleIndex(metric, vec, adj, deg, [vec.shape[0] - 1],
                         sph=(o, stored, spherical.LAYOUTS[layout]))
    mode = po.MODE_HYBRID if strategy == dk.Hybrid else po.MODE_PQ_ONLY
    r = ix.search_batch(qs, 10, L, bw, strategy)
    for i in range(qs.shape[0]):
        ids, ds, cm, hp = oix.search(qs[i], 10, L, bw, mode)
        assert list(r.ids[i, :len(ids)]) == list(ids), i
        assert r.distances[i, :len(ids)].tobytes() == ds.tobytes(), i
        assert (r.cmps[i], r.hops[i]) == (cm, hp), i

