"""Pin the oracle's k-means++ seeding (diskann-quantization/src/algorithms/kmeans/
plusplus.rs:238-498) and PQ train driver (diskann-quantization/src/product/train.rs:109-231)
given the uniform draws that stand in for the reference's RNG.

* Integer data keeps every distance and every f64 partial sum exact, so a brute-force
  restatement in exact integer arithmetic (int64 distances, Python-int cumulative sums)
  fixes which row each step picks: the first-row rule floor(u0 n), the min-distance update,
  the "first row with running sum >= threshold, distance > 0, not picked" rule, and the
  threshold u * scale of rand's UniformFloat over [0, total).
* n = K distinct rows: every row is picked exactly once.
* The failure cases: all-equal rows (zero total: InsufficientDiversity after one pick),
  K > n (DatasetTooSmall, every row picked), an infinite total (SawInfinity).
* The train driver equals its pinned pieces (k-means++, Lloyd assign / update) chunk by
  chunk, and on separated clusters its pivots are the cluster means (closed form).
The draws themselves (rand 0.9.4 / ChaCha12 in the reference) are parity unpinned."""
import math

import numpy as np

from oracle import pyoracle as po


def uniform_scale(s):
    if not (0.0 < s):
        return None
    sc = s
    max_rand = 1.0 - 2.0 ** -52
    while sc * max_rand >= s:
        sc = math.nextafter(sc, 0.0)
    return sc


def kpp_exact(x, K, u):
    """k-means++ on integer rows in exact arithmetic (valid while every f32 / f64 value
    involved is an exactly representable integer)."""
    x = x.astype(np.int64)
    n = x.shape[0]
    mind = [None] * n
    picked = set()
    i0 = min(int(u[0] * n), n - 1)
    chosen = [i0]
    picked.add(i0)
    for cur in range(1, min(K, n)):
        c = x[chosen[-1]]
        for r in range(n):
            d = int(((x[r] - c) ** 2).sum())
            mind[r] = d if mind[r] is None or d < mind[r] else mind[r]
        s = sum(mind)
        sc = uniform_scale(float(s))
        if sc is None:
            return chosen, po.KPP_INSUFFICIENT_DIVERSITY
        thr = u[cur] * sc
        run = 0
        pick = None
        for i in range(n):
            run += mind[i]
            if run >= thr and mind[i] > 0 and i not in picked:
                pick = i
                break
        if pick is None:
            return chosen, po.KPP_INSUFFICIENT_DIVERSITY
        chosen.append(pick)
        picked.add(pick)
    return chosen, (po.KPP_OK if len(chosen) == K else po.KPP_DATASET_TOO_SMALL)


def rows_of(centers, x):
    out = []
    for c in centers:
        hit = np.nonzero((x == c).all(axis=1))[0]
        out.append(int(hit[0]) if len(hit) else -1)
    return out


def test_integer_data_matches_exact_restatement():
    rng = np.random.default_rng(21)
    for n, ds, K in [(300, 2, 16), (257, 4, 32), (600, 3, 64), (50, 1, 20)]:
        x = rng.integers(-40, 41, (n, ds)).astype(np.float32)
        x = np.unique(x, axis=0)  # distinct rows: a picked centre names its row
        rng.shuffle(x)
        n = x.shape[0]
        for trial in range(3):
            u = rng.random(K)
            if trial == 0:
                u[1:4] = [0.0, 0.999999, 0.5]
            cen, sel, reason = po.kmeanspp(x, K, u)
            want, wreason = kpp_exact(x, K, u)
            assert reason == wreason
            assert sel == len(want)
            assert rows_of(cen[:sel], x) == want
            assert not cen[sel:].any()


def test_k_equals_n_picks_every_row():
    rng = np.random.default_rng(22)
    x = (rng.standard_normal((64, 4)) * 10).astype(np.float32)
    cen, sel, reason = po.kmeanspp(x, 64, rng.random(64))
    assert reason == po.KPP_OK and sel == 64
    assert sorted(rows_of(cen, x)) == list(range(64))


def test_failure_reasons():
    x = np.ones((40, 3), np.float32)
    cen, sel, reason = po.kmeanspp(x, 8, np.full(8, 0.3))
    assert (sel, reason) == (1, po.KPP_INSUFFICIENT_DIVERSITY)
    assert (cen[0] == 1).all() and not cen[1:].any()
    y = np.arange(12, dtype=np.float32).reshape(6, 2)
    cen, sel, reason = po.kmeanspp(y, 10, np.linspace(0, 0.9, 10))
    assert (sel, reason) == (6, po.KPP_DATASET_TOO_SMALL)
    assert sorted(rows_of(cen[:6], y)) == list(range(6))
    z = np.zeros((20, 2), np.float32)
    z[7] = [3e38, 3e38]  # its distance overflows f32: the f64 total is infinite
    cen, sel, reason = po.kmeanspp(z, 4, np.zeros(4))
    assert reason == po.KPP_SAW_INFINITY and sel == 1


def test_threshold_zero_picks_first_positive_unpicked_row():
    x = np.array([[0, 0], [0, 0], [1, 0], [5, 0], [9, 0]], np.float32)
    cen, sel, reason = po.kmeanspp(x, 3, np.array([0.0, 0.0, 0.0]))
    # first centre row 0; rows 0, 1 at distance 0; first positive row is 2
    assert rows_of(cen, x)[:2] == [0, 2]


def test_pq_train_equals_pinned_pieces():
    rng = np.random.default_rng(23)
    n, dim, M, K, iters = 700, 10, 3, 16, 4
    x = rng.standard_normal((n, dim)).astype(np.float32)
    off = po.chunk_offsets(dim, M)
    u = rng.random((M, K))
    piv = po.pq_train(x, off, K, iters, u)
    for m in range(M):
        xc = np.ascontiguousarray(x[:, off[m]:off[m + 1]])
        c, sel, reason = po.kmeanspp(xc, K, u[m])
        for _ in range(iters):
            nearest, _ = po.lloyd_assign(xc, c)
            c = po.lloyd_update(xc, nearest, K)
        assert piv[:, off[m]:off[m + 1]].tobytes() == c.tobytes()


def test_pq_train_separated_clusters_closed_form():
    rng = np.random.default_rng(24)
    K, per, dim, M = 8, 50, 8, 2
    centers = (np.arange(K, dtype=np.float32)[:, None] * 100.0) * np.ones((1, dim), np.float32)
    x = np.repeat(centers, per, axis=0) + rng.integers(-2, 3, (K * per, dim)).astype(np.float32)
    off = po.chunk_offsets(dim, M)
    u = rng.random((M, K))
    piv = po.pq_train(x, off, K, 3, u)
    for m in range(M):
        sl = slice(off[m], off[m + 1])
        means = np.stack([x[k * per:(k + 1) * per, sl].astype(np.float64).mean(0) for k in range(K)])
        got = piv[:, sl].astype(np.float64)
        # k-means++ seeds one centre per cluster here, so Lloyd converges to the means
        order = np.argsort(got[:, 0])
        assert np.allclose(got[order], means, atol=1e-5)
