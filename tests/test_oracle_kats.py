"""Pin the oracle against the reference's inline known-answer tests and against what the
mathematics fixes (FP64 bounds, closed forms, brute force)."""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as po

K = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))


def test_compute_pq_distance_kat():
    c = K["compute_pq_distance"]
    M = c["num_pq_chunks"]
    lut = np.arange(256 * M, dtype=np.float32).reshape(M, 256)
    data = np.array(c["pq_data"], dtype=np.uint8).reshape(c["n_pts"], M)
    for nid in c["neighbor_vector_ids"]:
        got = po.pq_adc(data[nid], lut)
        # the KAT's expected value: sequential f32 sum from 0.0 in chunk order
        exp = np.float32(0.0)
        for m in range(M):
            exp = np.float32(exp + lut[m, data[nid, m]])
        assert got == exp
        # and (exactly representable here) the integer closed form
        assert got == sum(m * 256 + int(data[nid, m]) for m in range(M))


@pytest.mark.parametrize("dim,m,exp", [tuple(x) for x in K["partition"]])
def test_partition_kat(dim, m, exp):
    assert list(po.chunk_offsets(dim, m)) == exp


def test_update_occlude_factor_kat():
    fmax = np.float32(3.4028234663852886e38)
    for dik in K["occlude_triangle_zero_jk"]:
        assert po.update_occlude_factor(po.L2, dik, 0.0, 1.0, 2.0) == fmax
    for alpha in [1.0, 1.1, 1.2, 1.3]:
        for dik, djk, cur, exp in K["occlude_triangle"]:
            assert po.update_occlude_factor(po.L2, dik, djk, cur, alpha) == exp
    for dik, djk, cur, alpha, exp in K["occlude_occluding"]:
        assert po.update_occlude_factor(po.INNER_PRODUCT, dik, djk, cur, alpha) == np.float32(exp)


def _l2_partials_reference(x, y):
    """Independent restatement of Strategy4x1 + f32x8 loads, written as the SIMD register
    program (4 accumulators of 8 lanes, epilogue vectors, masked tail, sum_tree)."""
    n = len(x)
    f = np.float32
    s = [[f(0)] * 8 for _ in range(4)]
    trip = n // 32
    epi = (n - 32 * trip) // 8

    def acc(reg, off):
        for l in range(8):
            c = f(x[off + l]) - f(y[off + l])
            reg[l] = f(np.float64(c) * np.float64(c) + np.float64(reg[l]))  # exact fma
    for i in range(trip):
        for j in range(4):
            acc(s[j], 32 * i + 8 * j)
    for j in range(epi):
        acc(s[j], 32 * trip + 8 * j)
    t = [f(f(s[0][l] + s[1][l]) + f(s[2][l] + s[3][l])) for l in range(8)]
    r = n % 8
    if r:
        base = n - r
        for l in range(8):
            a = f(x[base + l]) if l < r else f(0)
            b = f(y[base + l]) if l < r else f(0)
            c = a - b
            t[l] = f(np.float64(c) * np.float64(c) + np.float64(t[l]))
    a0, a1, a2, a3 = t[0] + t[4], t[1] + t[5], t[2] + t[6], t[3] + t[7]
    return f(f(a0 + a2) + f(a1 + a3))


@pytest.mark.parametrize("dim", [1, 3, 7, 8, 9, 16, 31, 32, 33, 40, 63, 96, 100, 128, 131])
def test_l2_order_matches_register_program(dim):
    # products of f32 values fit in f64 exactly, so f64 mul+add then rounding = fma
    rng = np.random.default_rng(dim)
    for _ in range(20):
        x = (rng.standard_normal(dim) * rng.choice([1e-3, 1.0, 1e3])).astype(np.float32)
        y = rng.standard_normal(dim).astype(np.float32)
        assert po.distance(po.L2, x, y) == _l2_partials_reference(x, y)


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE])
@pytest.mark.parametrize("dim", [4, 96, 128, 768])
def test_distance_fp64_bound(metric, dim):
    rng = np.random.default_rng(7 + dim)
    x = rng.standard_normal((50, dim)).astype(np.float32)
    y = rng.standard_normal((50, dim)).astype(np.float32)
    for a, b in zip(x, y):
        got = po.distance(metric, a, b)
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        if metric == po.L2:
            exp = np.sum((a64 - b64) ** 2)
            scale = exp
        elif metric == po.INNER_PRODUCT:
            exp = -np.dot(a64, b64)
            scale = np.dot(np.abs(a64), np.abs(b64))
        else:
            exp = 1 - np.dot(a64, b64) / np.linalg.norm(a64) / np.linalg.norm(b64)
            scale = 1.0
        assert abs(got - exp) <= 1e-5 * scale + 1e-6


def test_cosine_closed_forms():
    x = np.arange(1, 20, dtype=np.float32)
    assert abs(po.distance(po.COSINE, x, 2 * x)) <= 1e-6          # parallel
    assert po.distance(po.COSINE, x, np.zeros_like(x)) == 1.0    # zero norm guard
    e = np.zeros(12, np.float32)
    f = np.zeros(12, np.float32)
    e[0], f[1] = 1, 1
    assert po.distance(po.COSINE, e, f) == 1.0                    # orthogonal
    assert po.distance(po.COSINE, e, -e) == 2.0                   # opposite, clamped


def test_pq_lut_and_encode_brute_force():
    rng = np.random.default_rng(3)
    dim, M = 24, 5
    off = po.chunk_offsets(dim, M)
    piv = rng.standard_normal((256, dim)).astype(np.float32)
    x = rng.standard_normal((40, dim)).astype(np.float32)
    codes = po.pq_encode(piv, off, x)
    for i in range(x.shape[0]):
        for m in range(M):
            s, e = off[m], off[m + 1]
            d = ((x[i, s:e].astype(np.float64) - piv[:, s:e]) ** 2).sum(1)
            assert d[codes[i, m]] <= d.min() + 1e-4
    lut = po.pq_lut(po.L2, piv, off, x[0])
    for m in range(M):
        s, e = off[m], off[m + 1]
        exp = ((x[0, s:e].astype(np.float64) - piv[:, s:e]) ** 2).sum(1)
        np.testing.assert_allclose(lut[m], exp, rtol=1e-5, atol=1e-6)
        # each LUT entry is exactly the chunk's SquaredL2 in the reference order
        for c in range(0, 256, 37):
            assert lut[m, c] == po.distance(po.L2, x[0, s:e], piv[c, s:e])
    ip = po.pq_lut(po.INNER_PRODUCT, piv, off, x[0])
    assert ip[0, 7] == po.distance(po.INNER_PRODUCT, x[0, off[0]:off[1]], piv[7, off[0]:off[1]])


def test_search_recall_and_invariants_small():
    rng = np.random.default_rng(11)
    n, dim = 1500, 16
    x = rng.standard_normal((n, dim)).astype(np.float32)
    med, mean = po.medoid(x)
    vec = np.vstack([x, x[med:med + 1]])
    adj, deg = po.build_graph(po.L2, vec, R=16, R_max=20, L_build=40, alpha=1.2,
                              starts=[n], ids=np.arange(n), batch=64)
    # invariants: degree bound, no self loops, no duplicates
    for i in range(n):
        lst = adj[i, :deg[i]]
        assert deg[i] <= 20 and i not in lst and len(set(lst.tolist())) == len(lst)
    q = rng.standard_normal((50, dim)).astype(np.float32)
    ix = po.OracleIndex(po.L2, vec, adj, deg, [n])
    gt, _ = po.bruteforce_f64(po.L2, x, q, 10)
    ids, ds, cm, hp, _, _ = ix.search_batch(q, 10, 64, nthreads=4)
    recall = np.mean([len(set(ids[i]) & set(gt[i])) / 10 for i in range(len(q))])
    assert recall > 0.9
    # results sorted, distances equal to direct evaluation
    for i in range(len(q)):
        assert np.all(np.diff(ds[i]) >= 0)
        assert ds[i, 0] == po.distance(po.L2, q[i], x[ids[i, 0]])


def test_lloyd_closed_form_separated_clusters():
    rng = np.random.default_rng(5)
    centers = np.array([[0, 0], [100, 0], [0, 100]], np.float32)
    x = np.vstack([c + rng.standard_normal((30, 2)).astype(np.float32) for c in centers])
    lab, _ = po.lloyd_assign(x, centers)
    assert (lab == np.repeat([0, 1, 2], 30)).all()
    c2 = po.lloyd_update(x, lab, 4)  # cluster 3 empty -> zero vector
    np.testing.assert_allclose(c2[:3], [x[i * 30:(i + 1) * 30].astype(np.float64).mean(0)
                                        for i in range(3)], rtol=1e-6)
    assert (c2[3] == 0).all()
