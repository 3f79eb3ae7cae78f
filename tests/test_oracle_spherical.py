"""1-bit spherical quantization on the oracle (diskann-quantization/src/spherical).

Pins:
* the reference's own golden baselines (test/generated/spherical/1bit_*_double_hadamard.json,
  committed as tests/golden/spherical_1bit.json with the 16 x 8 dataset of
  spherical/iface.rs test_dataset): the centroid and mean norm of training, every
  compressed byte (bits and f16 metadata), every data-data distance and every query
  distance for the same-as-data, 4-bit transposed and full-precision layouts, with and
  without the inner-product rescale - all compared as f32, bit for bit;
* the Hadamard transform against the Sylvester matrix in f64 for lengths 1 .. 4096, and
  exactly on integer inputs (where every order of additions is exact);
* f32 -> f16 rounding against numpy's float16 conversion (round to nearest even,
  subnormals, overflow);
* transforms preserve norms (orthogonal) and the padding / subsample variants keep their
  documented shapes.
The 64-element micro-kernel's order of additions (lengths >= 64) has no golden vector:
parity for those lengths rests on the restatement plus the f64 bound (parity partial)."""
import numpy as np
import pytest

import sphutil
from oracle import pyoracle as po

G = sphutil.load()
X = np.array(G["dataset"], np.float32)


def quantizer(metric_name):
    b = G["baselines"][metric_name]
    q = sphutil.parse_quantizer(b["serialized_quantizer"])
    assert q["kind"] == "double_hadamard" and q["nbits"] == 1 and q["subsample"] is None
    sq = po.Spherical(sphutil.METRICS[metric_name], q["centroid"], q["mean_norm"],
                      q["pre_scale"], po.SPH_DOUBLE_HADAMARD, q["signs0"], q["signs1"])
    return sq, q, b


@pytest.mark.parametrize("metric", sorted(sphutil.METRICS))
def test_training_statistics_match_golden(metric):
    _, q, _ = quantizer(metric)
    c, mn = po.sph_train_stats(sphutil.METRICS[metric], X)
    assert np.array_equal(c, q["centroid"])
    assert np.float32(mn) == np.float32(q["mean_norm"])


@pytest.mark.parametrize("metric", sorted(sphutil.METRICS))
def test_compressed_bytes_match_golden(metric):
    sq, _, b = quantizer(metric)
    assert np.array_equal(sq.compress(X), np.array(b["compressed_vectors"], np.uint8))


@pytest.mark.parametrize("metric", sorted(sphutil.METRICS))
def test_data_distances_match_golden(metric):
    sq, _, b = quantizer(metric)
    c = sq.compress(X)
    k = 0
    for i in range(16):
        for j in range(i, 16):
            assert np.float32(sq.dist_data(c[i], c[j])) == np.float32(b["data_distances"][k])
            k += 1


def layout_distances(sq, layout, x, codes):
    if layout == "same_as_data":
        qc = sq.compress(x)[0]
        return [sq.dist_data(qc, codes[j]) for j in range(len(codes))]
    if layout == "four_bit_transposed":
        c, m = sq.compress_query4(x)
        return [sq.dist_query4(c, m, codes[j]) for j in range(len(codes))]
    v, m = sq.compress_full(x)
    return [sq.dist_full(v, m, codes[j]) for j in range(len(codes))]


@pytest.mark.parametrize("metric", sorted(sphutil.METRICS))
@pytest.mark.parametrize("rescaled", [False, True])
def test_query_distances_match_golden(metric, rescaled):
    sq, _, b = quantizer(metric)
    key = "rescaled_query_distances" if rescaled else "query_distances"
    if key not in b:
        pytest.skip("rescaled distances exist for inner product only")
    codes = sq.compress(X)
    assert sorted(L["layout"] for L in b[key]) == ["four_bit_transposed", "full_precision",
                                                  "same_as_data"]
    for L in b[key]:
        for qi in range(16):
            x = sq.rescale(X[qi]) if rescaled else X[qi]
            got = np.array(layout_distances(sq, L["layout"], x, codes), np.float32)
            assert np.array_equal(got, np.array(L["distances"][qi], np.float32)), \
                (L["layout"], qi)


def sylvester(n):
    h = np.array([[1.0]])
    while h.shape[0] < n:
        h = np.block([[h, h], [h, -h]])
    return h


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 256, 1024, 4096])
def test_hadamard_against_matrix(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(np.float32)
    got = po.hadamard(x)
    exp = sylvester(n) @ x.astype(np.float64) / np.sqrt(n)
    assert np.abs(got - exp).max() <= 1e-5 * max(1.0, np.abs(exp).max())
    xi = rng.integers(-50, 50, n).astype(np.float32)
    got = po.hadamard(xi)
    hx = (sylvester(n) @ xi.astype(np.float64)).astype(np.float32)   # exact integer sums
    m = np.float32(1.0) / np.sqrt(np.float32(n)) if n > 1 else np.float32(1.0)
    assert np.array_equal(got, hx * m)


def test_f16_rounding_matches_numpy():
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-9, 6, 20000),
        np.array([0.0, -0.0, 65504.0, 65519.0, 65520.0, 1e9, -1e9, 6e-8, 3e-8, 2.98e-8,
                  5.96e-8, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11], np.float32)])
    with np.errstate(over="ignore"):
        exp = vals.astype(np.float16).view(np.uint16)
    for v, e in zip(vals, exp):
        assert po.f32_to_f16_bits(v) == int(e), v


@pytest.mark.parametrize("kind,dim", [(po.SPH_DOUBLE_HADAMARD, 100), (po.SPH_DOUBLE_HADAMARD, 128),
                                      (po.SPH_PADDING_HADAMARD, 100), (po.SPH_NULL, 37)])
def test_transforms_preserve_norms(kind, dim):
    rng = np.random.default_rng(dim + kind)
    s0 = rng.integers(0, 2, dim).astype(bool)
    s1 = rng.integers(0, 2, dim).astype(bool)
    sq = po.Spherical(po.L2, np.zeros(dim, np.float32), 1.0, 1.0, kind, s0, s1,
                      padded_dim=128)
    x = rng.standard_normal(dim).astype(np.float32)
    t = sq.transform(x)
    assert t.shape[0] == (128 if kind == po.SPH_PADDING_HADAMARD else dim)
    assert abs(np.linalg.norm(t) - np.linalg.norm(x)) < 1e-5 * np.linalg.norm(x)
    if kind == po.SPH_NULL:
        assert np.array_equal(t, x)


def test_subsample_rescales():
    dim = 64
    rng = np.random.default_rng(2)
    s0 = rng.integers(0, 2, dim).astype(bool)
    s1 = rng.integers(0, 2, dim).astype(bool)
    sub = np.sort(rng.choice(dim, 40, replace=False))
    full = po.Spherical(po.L2, np.zeros(dim, np.float32), 1.0, 1.0, po.SPH_DOUBLE_HADAMARD, s0, s1)
    part = po.Spherical(po.L2, np.zeros(dim, np.float32), 1.0, 1.0, po.SPH_DOUBLE_HADAMARD, s0, s1,
                        subsample=sub)
    x = rng.standard_normal(dim).astype(np.float32)
    assert np.array_equal(part.transform(x), full.transform(x)[sub] * np.sqrt(np.float32(64 / 40)))


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE])
def test_distances_track_full_precision(metric):
    """Sanity of the estimator on real data (the reference's providers test only checks a
    relative error bound): correlation of 1-bit estimates with exact distances."""
    rng = np.random.default_rng(7)
    dim = 128
    x = rng.standard_normal((300, dim)).astype(np.float32)
    c, mn = po.sph_train_stats(metric, x)
    sq = po.Spherical(metric, c, mn, 1.0, po.SPH_DOUBLE_HADAMARD,
                      rng.integers(0, 2, dim).astype(bool), rng.integers(0, 2, dim).astype(bool))
    codes = sq.compress(x)
    q = x[0]
    cq, mq = sq.compress_query4(q)
    est = np.array([sq.dist_query4(cq, mq, codes[j]) for j in range(1, 300)])
    ex = np.array([po.distance(metric, q, x[j]) for j in range(1, 300)])
    assert np.corrcoef(est, ex)[0, 1] > 0.6


def sph_graph_index(metric, layout, n=300, dim=40, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n + 1, dim)).astype(np.float32)
    if metric != po.L2:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    x[n] = x[:n].mean(axis=0)
    adj, deg = po.build_graph(metric, x, R=16, R_max=20, L_build=40, alpha=1.2,
                              starts=[n], ids=np.arange(n))
    c, mn = po.sph_train_stats(metric, x[:n])
    sq = po.Spherical(metric, c, mn, 1.0, po.SPH_DOUBLE_HADAMARD,
                      rng.integers(0, 2, dim).astype(bool), rng.integers(0, 2, dim).astype(bool))
    codes = sq.compress(x)
    ix = po.OracleIndex(metric, x, adj, deg, [n], sph=(sq, codes, layout))
    return ix, x, sq, codes


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE])
@pytest.mark.parametrize("layout", [0, 1])
def test_traversal_exhaustive_is_bruteforce(metric, layout):
    """With L covering the graph, the quantized search returns the brute-force top-k of
    the layout's 1-bit distances (inner-product queries rescaled as the search strategy
    asks)."""
    ix, x, sq, codes = sph_graph_index(metric, layout)
    n = x.shape[0] - 1
    seen, todo = {n}, [n]
    while todo:
        u = todo.pop()
        for v in ix.adj[u, :ix.deg[u]]:
            if int(v) not in seen:
                seen.add(int(v))
                todo.append(int(v))
    reach = np.array(sorted(seen - {n}))
    rng = np.random.default_rng(11)
    for _ in range(4):
        q = rng.standard_normal(x.shape[1]).astype(np.float32)
        ids, ds, cm, hp = ix.search(q, 10, n + 1, mode=po.MODE_PQ_ONLY)
        qq = sq.rescale(q) if metric == po.INNER_PRODUCT else q
        if layout == 0:
            qc = sq.compress(qq)[0]
            bd = np.array([sq.dist_data(qc, codes[r]) for r in reach], np.float32)
        else:
            c4, m4 = sq.compress_query4(qq)
            bd = np.array([sq.dist_query4(c4, m4, codes[r]) for r in reach], np.float32)
        order = np.argsort(bd, kind="stable")[:10]
        assert np.array_equal(np.sort(ds), np.sort(bd[order]))
