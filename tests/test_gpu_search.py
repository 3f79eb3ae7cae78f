"""GPU parity: the HIP beam search / PQ / distance kernels (through the C ABI) against
the oracle on the same inputs. Integer outputs (ids, PQ codes, cmps, hops) must be
bit-exact; f32 distances are compared bit-for-bit as well (the kernels reproduce the
reference's reduction order), which is stricter than the 1e-5 relative tolerance the
north star allows."""
import json
import os

import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd import quantization as q
from gridutil import grid
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "grid_cases.json")))


def log(msg):
    print(msg, flush=True)


def make_index(metric, vec, adj, deg, nstart=1, R=None):
    n = vec.shape[0] - nstart
    R_max = adj.shape[1]
    ix = dk.DiskANNIndex(metric, vec.shape[1], n, R or R_max, R_max, num_start=nstart)
    ix.set_vectors(vec[:n])
    ix.set_start_points(vec[n:])
    ix.set_graph(adj, deg)
    return ix


@pytest.mark.parametrize("case", sorted(FIX["grid_search"]))
def test_grid_search_golden_gpu(case):
    log("grid golden %s" % case)
    c = FIX["grid_search"][case]
    vec, adj, deg, starts = grid(c["grid_dims"], c["grid_size"])
    ix = make_index(dk.Metric.L2, vec, adj, deg)
    for s in c["searches"]:
        ids, ds, cmps, hops = ix.search(dk.Knn(c["k"], c["L"], s["beam_width"]),
                                        dk.FullPrecision, np.array(s["query"], np.float32))
        assert [int(i) for i in ids] == [e[0] for e in s["results"]]
        assert [float(d) for d in ds] == [e[1] for e in s["results"]]
        assert cmps == s["comparisons"] and hops == s["hops"]


_DATA = {}


def dataset(metric, n=3000, dim=128, seed=0, R=24, L_build=48):
    key = (metric, n, dim, seed, R, L_build)
    if key in _DATA:
        return _DATA[key]
    rng = np.random.default_rng(seed)
    centers = rng.standard_normal((16, dim)).astype(np.float32) * 3
    x = (centers[rng.integers(0, 16, n)] + rng.standard_normal((n, dim))).astype(np.float32)
    if metric in (po.COSINE, po.INNER_PRODUCT):
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    med, _ = po.medoid(x)
    vec = np.vstack([x, x[med:med + 1]])
    R_max = int(1.3 * R)
    adj, deg = po.build_graph(metric, vec, R=R, R_max=R_max, L_build=L_build, alpha=1.2,
                              starts=[n], ids=np.arange(n), batch=128)
    qs = (centers[rng.integers(0, 16, 64)] + rng.standard_normal((64, dim))).astype(np.float32)
    if metric in (po.COSINE, po.INNER_PRODUCT):
        qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    _DATA[key] = (vec, adj, deg, qs)
    return _DATA[key]


@pytest.mark.parametrize("metric,dim", [(po.L2, 128), (po.L2, 96), (po.L2, 37),
                                        (po.INNER_PRODUCT, 64), (po.COSINE, 48)])
@pytest.mark.parametrize("L,bw", [(10, 1), (50, 1), (64, 2), (100, 4)])
def test_fp_search_matches_oracle(metric, dim, L, bw):
    log("fp search metric=%d dim=%d L=%d bw=%d" % (metric, dim, L, bw))
    vec, adj, deg, qs = dataset(metric, dim=dim)
    ix = make_index(metric, vec, adj, deg)
    oix = po.OracleIndex(metric, vec, adj, deg, [vec.shape[0] - 1])
    k = 10
    r = ix.search_batch(qs, k, L, bw, dk.FullPrecision)
    for i in range(qs.shape[0]):
        ids, ds, cm, hp = oix.search(qs[i], k, L, bw)
        assert r.counts[i] == len(ids)
        assert list(r.ids[i, :len(ids)]) == list(ids), i
        assert r.distances[i, :len(ids)].tobytes() == ds.tobytes(), i
        assert (r.cmps[i], r.hops[i]) == (cm, hp), i


def pq_setup(vec, M, seed=1):
    rng = np.random.default_rng(seed)
    dim = vec.shape[1]
    off = po.chunk_offsets(dim, M)
    piv = vec[rng.choice(vec.shape[0], 256, replace=False)] + \
        rng.standard_normal((256, dim)).astype(np.float32) * 0.05
    return piv.astype(np.float32), off


@pytest.mark.parametrize("metric,dim,M", [(po.L2, 128, 32), (po.L2, 96, 48), (po.L2, 40, 7),
                                          (po.INNER_PRODUCT, 64, 16)])
def test_pq_encode_lut_match_oracle(metric, dim, M):
    log("pq encode/lut dim=%d M=%d" % (dim, M))
    vec, adj, deg, qs = dataset(metric, dim=dim)
    piv, off = pq_setup(vec, M)
    codes_g = q.pq_encode(piv, off, vec)
    codes_o = po.pq_encode(piv, off, vec)
    assert np.array_equal(codes_g, codes_o)
    lut_g = q.pq_lut(metric, piv, off, qs[:8])
    for i in range(8):
        lut_o = po.pq_lut(metric, piv, off, qs[i])
        assert lut_g[i].tobytes() == lut_o.tobytes()


@pytest.mark.parametrize("metric,dim,M", [(po.L2, 128, 32), (po.L2, 96, 48), (po.L2, 40, 7),
                                          (po.INNER_PRODUCT, 64, 16)])
@pytest.mark.parametrize("strategy", [dk.Hybrid, dk.Quantized])
@pytest.mark.parametrize("L,bw", [(32, 1), (64, 2)])
@pytest.mark.parametrize("adc_mode", [0, 1])
def test_pq_search_matches_oracle(metric, dim, M, strategy, L, bw, adc_mode):
    log("pq search metric=%d dim=%d M=%d strat=%d L=%d adc=%d" % (metric, dim, M, strategy, L,
                                                                  adc_mode))
    vec, adj, deg, qs = dataset(metric, dim=dim)
    piv, off = pq_setup(vec, M)
    ix = make_index(metric, vec, adj, deg)
    ix.set_tuning(0, adc_mode)
    ix.set_pq(piv, off)
    ix.encode()
    codes = ix.get_codes()
    assert np.array_equal(codes, po.pq_encode(piv, off, vec))
    oix = po.OracleIndex(metric, vec, adj, deg, [vec.shape[0] - 1], codes, piv, off)
    mode = po.MODE_HYBRID if strategy == dk.Hybrid else po.MODE_PQ_ONLY
    r = ix.search_batch(qs, 10, L, bw, strategy)
    for i in range(qs.shape[0]):
        ids, ds, cm, hp = oix.search(qs[i], 10, L, bw, mode)
        assert list(r.ids[i, :len(ids)]) == list(ids), i
        assert r.distances[i, :len(ids)].tobytes() == ds.tobytes(), i
        assert (r.cmps[i], r.hops[i]) == (cm, hp), i


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE])
@pytest.mark.parametrize("dim", [1, 5, 8, 17, 32, 33, 96, 100, 128, 768, 1536])
def test_distance_batch_bit_exact(metric, dim):
    rng = np.random.default_rng(dim)
    base = rng.standard_normal((50, dim)).astype(np.float32)
    qs = rng.standard_normal((4, dim)).astype(np.float32)
    ids = rng.integers(0, 50, (4, 9)).astype(np.uint32)
    out = q.distance_batch(metric, qs, base, ids)
    for i in range(4):
        for j in range(9):
            assert out[i, j] == po.distance(metric, qs[i], base[ids[i, j]]), (i, j)


def test_mixed_k_and_L_batch_equals_alone():
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    rng = np.random.default_rng(3)
    ks = rng.choice([1, 5, 10], qs.shape[0]).astype(np.uint32)
    Ls = np.maximum(ks, rng.integers(8, 120, qs.shape[0])).astype(np.uint32)
    r = ix.search_batch(qs, ks, Ls, 1, dk.FullPrecision)
    for i in range(0, qs.shape[0], 7):
        ids, ds, cm, hp = ix.search(dk.Knn(int(ks[i]), int(Ls[i])), dk.FullPrecision, qs[i])
        assert list(r.ids[i, :r.counts[i]]) == list(ids)
        assert r.distances[i, :r.counts[i]].tobytes() == ds.tobytes()


def test_visited_spill_path_matches_oracle():
    # a tiny LDS visited table forces the global (HBM) spill table into use
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    ix.set_tuning(visited_lds_slots=256)
    assert ix.tuning == (256, 1)
    oix = po.OracleIndex(po.L2, vec, adj, deg, [vec.shape[0] - 1])
    r = ix.search_batch(qs, 10, 100, 1, dk.FullPrecision)
    for i in range(qs.shape[0]):
        ids, ds, cm, hp = oix.search(qs[i], 10, 100, 1)
        assert list(r.ids[i, :len(ids)]) == list(ids)
        assert (r.cmps[i], r.hops[i]) == (cm, hp)


def test_knn_errors_like_reference():
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    with pytest.raises(dk.KnnSearchError):
        ix.search_batch(qs[:2], [10, 100], [50, 32])   # L < k on the second query
    with pytest.raises(dk.KnnSearchError):
        ix.search_batch(qs[:2], 0, 10)
    with pytest.raises(dk.KnnSearchError):
        ix.search_batch(qs[:2], 5, 10, beam_width=0)
