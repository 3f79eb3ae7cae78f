"""Scalar quantization (8 bits) on the oracle: training, compression with compensation,
the compensated distances and the scalar-quantized traversal.

Pins (none of them retypes the oracle's formulas):
* the reference's documented example (diskann-quantization/src/scalar/quantizer.rs:20-80):
  rows of -1 and +1 with one standard deviation give inner product 5 and squared L2 20;
* training on data whose f64 sums are exact in any order (integer rows, eight of them):
  scale = 2 * stdevs * max std, shift = mean - stdevs * max std, exactly as f32, and the
  mean norm against a sum taken in row order (train.rs test_train);
* compression against the reference's own coverage test (quantizer.rs:540-760): values
  below the shift code 0, above the range code 255, the rest reconstruct within half a
  step; every code occurs; the compensation equals the fma sum of code times shift;
  exact half steps round away from zero;
* distances against the reconstructed vectors X = (scale / 255) X' + shift in f64
  (vectors.rs test_compensated_distance): a dropped compensation or shift term fails;
* traversal: with L covering the whole graph the quantized search returns the
  brute-force top-k of the scalar-quantized distances."""
import numpy as np
import pytest

from oracle import pyoracle as po


def test_documented_example():
    x = np.array([[-1.0] * 5, [1.0] * 5], np.float32)
    scale, shift, mean_norm = po.sq_train(x, 1.0)
    codes, comp = po.sq_compress(scale, shift, x)
    ssn = po.sq_shift_square_norm(shift)
    ip = po.sq_distance(po.INNER_PRODUCT, scale, ssn, codes[0], comp[0], codes[1], comp[1])
    l2 = po.sq_distance(po.L2, scale, ssn, codes[0], comp[0], codes[1], comp[1])
    assert abs(ip - 5.0) < 1e-5
    assert abs(l2 - 20.0) < 1e-5


@pytest.mark.parametrize("stdevs", [1.0, 1.5, 2.0])
def test_train_exact_statistics(stdevs):
    rng = np.random.default_rng(4)
    x = rng.integers(-20, 21, (8, 12)).astype(np.float32)
    scale, shift, mean_norm = po.sq_train(x, stdevs)
    xd = x.astype(np.float64)
    mean = xd.sum(axis=0) / 8.0            # exact: small integers, n = 8
    var = ((xd - mean) ** 2).sum(axis=0) / 8.0
    p = np.sqrt(var.max()) * stdevs
    assert np.float32(scale) == np.float32(2.0 * p)
    assert np.array_equal(shift, (mean - p).astype(np.float32))
    acc = 0.0
    for r in range(8):
        acc += float(np.sqrt(np.sum(xd[r] ** 2)))   # exact squares, row-order norm sum
    assert np.float32(mean_norm) == np.float32(acc / 8.0)


def test_train_real_data_close_to_float64():
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((500, 40)) * rng.uniform(0.5, 3, 40) + rng.uniform(-5, 5, 40))
    x = x.astype(np.float32)
    scale, shift, mean_norm = po.sq_train(x, 2.0)
    xd = x.astype(np.float64)
    p = 2.0 * np.sqrt(xd.var(axis=0).max())
    assert abs(scale - 2 * p) <= 1e-6 * 2 * p
    assert np.allclose(shift, xd.mean(axis=0) - p, rtol=0, atol=1e-5)
    assert abs(mean_norm - np.linalg.norm(xd, axis=1).mean()) < 1e-4


def test_compress_coverage():
    rng = np.random.default_rng(6)
    dim, step = 16, 2.0
    shift = rng.integers(-10, 11, dim).astype(np.float32)
    scale = np.float32(step * 255.0)
    assert po.sq_shift_square_norm(shift) == np.float32((shift.astype(np.float64) ** 2).sum())
    lo, hi = -3 * step, 255 * step + 3 * step
    base = np.arange(lo, hi + step / 3, step / 3.1, dtype=np.float32)
    offs = np.stack([rng.permutation(base) for _ in range(dim)], axis=1)
    seen = [set() for _ in range(dim)]
    for o in offs:
        q = (shift + o).astype(np.float32)
        codes, comp = po.sq_compress(scale, shift, q)
        c = codes[0]
        acc = np.float32(0.0)
        for d in range(dim):
            acc = np.float32(np.float64(c[d]) * np.float64(shift[d]) + np.float64(acc))
            seen[d].add(int(c[d]))
            if o[d] <= 0:
                assert c[d] == 0
            elif o[d] >= 255 * step:
                assert c[d] == 255
            else:
                assert abs(o[d] - scale * np.float32(c[d]) * np.float32(1 / 255)) <= step / 2 + 1e-4
        assert comp[0] == np.float32(scale * np.float32(1.0 / 255.0)) * acc
    assert all(len(s) == 256 for s in seen)


def test_compress_rounds_half_away_and_rejects_nan():
    shift = np.zeros(4, np.float32)
    codes, _ = po.sq_compress(np.float32(255.0), shift,
                              np.array([0.5, 1.5, 2.49, 254.5], np.float32))
    assert list(codes[0]) == [1, 2, 2, 255]
    with pytest.raises(ValueError):
        po.sq_compress(np.float32(255.0), shift, np.array([0, np.nan, 1, 2], np.float32))


@pytest.mark.parametrize("dim", [8, 37, 128])
def test_distances_match_reconstruction(dim):
    rng = np.random.default_rng(dim)
    x = (rng.standard_normal((60, dim)) + rng.uniform(-2, 2, dim)).astype(np.float32)
    scale, shift, mean_norm = po.sq_train(x, 2.0)
    codes, comp = po.sq_compress(scale, shift, x)
    ssn = po.sq_shift_square_norm(shift)
    a = np.float64(scale) / 255.0
    rec = a * codes.astype(np.float64) + shift.astype(np.float64)
    for i, j in [(0, 1), (2, 7), (10, 59), (5, 5)]:
        l2 = po.sq_distance(po.L2, scale, ssn, codes[i], comp[i], codes[j], comp[j])
        ip = po.sq_distance(po.INNER_PRODUCT, scale, ssn, codes[i], comp[i], codes[j], comp[j])
        cn = po.sq_distance(po.COSINE_NORMALIZED, scale, ssn, codes[i], comp[i], codes[j],
                            comp[j])
        e_l2 = np.sum((rec[i] - rec[j]) ** 2)
        e_ip = -np.dot(rec[i], rec[j])
        tol = 1e-5 * (np.dot(rec[i], rec[i]) + np.dot(rec[j], rec[j]) + 1.0)
        assert abs(l2 - e_l2) <= tol, (l2, e_l2)
        assert abs(ip - e_ip) <= tol, (ip, e_ip)
        assert abs(cn - e_l2 / 2) <= tol, (cn, e_l2 / 2)


def test_rescale():
    x = np.array([3.0, 4.0, 0.0], np.float32)
    assert np.allclose(po.sq_rescale(x, 10.0), [6.0, 8.0, 0.0])
    assert np.array_equal(po.sq_rescale(np.zeros(3, np.float32), 2.0), np.zeros(3))


def small_graph_index(metric, n=300, dim=24, seed=3):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((n + 1, dim))).astype(np.float32)
    if metric != po.L2:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    x[n] = x[:n].mean(axis=0)
    adj, deg = po.build_graph(metric, x, R=16, R_max=20, L_build=40, alpha=1.2,
                              starts=[n], ids=np.arange(n))
    scale, shift, mean_norm = po.sq_train(x[:n], 2.0)
    codes, comp = po.sq_compress(scale, shift, x)
    ix = po.OracleIndex(metric, x, adj, deg, [n], sq=(scale, shift, mean_norm, codes, comp))
    return ix, x, (scale, shift, mean_norm, codes, comp)


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE_NORMALIZED])
def test_quantized_traversal_exhaustive_is_bruteforce(metric):
    ix, x, (scale, shift, mean_norm, codes, comp) = small_graph_index(metric)
    n = x.shape[0] - 1
    ssn = po.sq_shift_square_norm(shift)
    seen, todo = {n}, [n]           # rows reachable from the start point
    while todo:
        u = todo.pop()
        for v in ix.adj[u, :ix.deg[u]]:
            if int(v) not in seen:
                seen.add(int(v))
                todo.append(int(v))
    reach = np.array(sorted(seen - {n}))
    rng = np.random.default_rng(11)
    for _ in range(5):
        q = rng.standard_normal(x.shape[1]).astype(np.float32)
        ids, ds, cm, hp = ix.search(q, 10, n + 1, mode=po.MODE_PQ_ONLY)
        qq = po.sq_rescale(q, mean_norm) if metric == po.INNER_PRODUCT else q
        qc, qcomp = po.sq_compress(scale, shift, qq)
        bd = np.array([po.sq_distance(metric, scale, ssn, qc[0], qcomp[0], codes[r], comp[r])
                       for r in reach], np.float32)
        order = np.argsort(bd, kind="stable")[:10]
        assert np.array_equal(np.sort(ds), np.sort(bd[order]))
        assert set(ids.tolist()) <= set(reach[bd <= bd[order[-1]]].tolist())


def test_quantized_rerank_returns_exact_distances():
    ix, x, _ = small_graph_index(po.L2)
    q = x[5] + 0.01
    ids, ds, cm, hp = ix.search(q, 10, 60, mode=po.MODE_HYBRID)
    exact = np.array([po.distance(po.L2, q, x[i]) for i in ids], np.float32)
    assert np.array_equal(ds, exact) and np.all(np.diff(ds) >= 0)
