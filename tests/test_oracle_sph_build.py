"""Pin the oracle's 1-bit spherical build (Quantized::build() over a SphericalStore:
diskann-providers/src/model/graph/provider/async_/inmem/spherical.rs:293-350, 515-533,
605-650) by the identities its two distance sources must satisfy:

* the insert search compresses its query in the data layout (SameAsData) without the
  inner-product rescale (is_search = false), so a data row searching for itself scores every
  node with exactly the data x data distance of the two stored codes (the prune's distance,
  Spherical.dist_data, itself pinned bit for bit by the reference's golden baselines); with
  the rescale of the search strategies the inner-product scores differ;
* robust_prune_list under the spherical strategy equals occlude_list on the pool the
  code distances give;
* the spherical graph differs from the full-precision one (the strategy is taken)."""
import numpy as np
import pytest

from oracle import pyoracle as po


def setup(metric, n=400, dim=64, seed=2, normalize=True):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((8, dim)).astype(np.float32) * 3
    x = (c[rng.integers(0, 8, n)] + rng.standard_normal((n, dim))).astype(np.float32)
    if metric != po.L2 and normalize:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    med, _ = po.medoid(x)
    vec = np.vstack([x, x[med:med + 1]])
    cen, mn = po.sph_train_stats(metric, vec[:-1])
    q = po.Spherical(metric, cen, mn, 1.0, po.SPH_NULL)
    return vec, q, q.compress(vec)


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE])
def test_insert_search_scores_are_code_pair_distances(metric):
    # inner product on rows of uneven norm: there the rescale of the search strategies shows
    vec, q, codes = setup(metric, normalize=metric != po.INNER_PRODUCT)
    n = vec.shape[0] - 1
    adj, deg = po.build_graph(metric, vec, R=12, R_max=16, L_build=32, alpha=1.2, starts=[n],
                              ids=np.arange(n), batch=1)
    oix = po.OracleIndex(metric, vec, adj, deg, [n], sph=(q, codes, 0, True))
    for i in range(0, n, 40):
        ids, ds, _, _ = oix.search(vec[i], 10, 32, 1, po.MODE_PQ_ONLY)
        for j, d in zip(ids, ds):
            assert np.float32(d).tobytes() == np.float32(q.dist_data(codes[i], codes[j])).tobytes()
    if metric == po.INNER_PRODUCT:
        rix = po.OracleIndex(metric, vec, adj, deg, [n], sph=(q, codes, 0))
        diff = 0
        for i in range(0, n, 40):
            ids, ds, _, _ = rix.search(vec[i], 10, 32, 1, po.MODE_PQ_ONLY)
            diff += sum(d != q.dist_data(codes[i], codes[j]) for j, d in zip(ids, ds))
        assert diff > 0


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT])
def test_prune_list_is_occlusion_over_code_distances(metric):
    vec, q, codes = setup(metric, seed=4)
    rng = np.random.default_rng(1)
    R = 10
    for loc in range(0, 400, 57):
        lst = rng.choice(400, 100, replace=False).astype(np.uint32)
        lst = lst[lst != loc]
        got = po.prune_list(metric, vec, loc, lst, R, 1.2, sph=(q, codes))
        # the pool: code distances from the location, stable sort, then occlusion with code
        # distances between candidates: run the same occlusion on surrogate vectors is not
        # possible, so compare with a direct restatement of occlude_list (index.rs:2570-2782)
        d = np.array([q.dist_data(codes[loc], codes[j]) for j in lst], np.float32)
        order = np.argsort(d, kind="stable")
        pool, pd = lst[order], d[order]
        want = occlude(metric, pool, pd, lambda a, b: q.dist_data(codes[a], codes[b]), R, 1.2)
        assert list(got) == want


def occlude(metric, pool, pd, dist, R, alpha):
    """occlude_list without saturation (diskann/src/graph/index.rs:2570-2782) on a sorted
    pool, in plain Python."""
    n = len(pool)
    occ = [0.0] * n
    acc = []
    cur = 1.0
    inc = min(alpha, 1.2)
    fmax = float(np.finfo(np.float32).max)
    while len(acc) < R:
        for i in range(n):
            if len(acc) >= R:
                break
            if occ[i] > cur:
                continue
            of = occ[i]
            for j in acc:
                if j >= i:      # only accepted entries earlier in the pool occlude
                    continue
                dj = dist(int(pool[j]), int(pool[i]))
                of = po.update_occlude_factor(metric, float(pd[i]), float(np.float32(dj)), of, cur)
                if of > cur:
                    break
            if of > cur:
                occ[i] = of
                continue
            occ[i] = fmax
            acc.append(i)
        if cur == alpha:
            break
        cur = float(np.float32(min(np.float32(cur) * np.float32(inc), np.float32(alpha))))
    return [int(pool[i]) for i in acc]


def test_sph_build_differs_from_full_precision():
    vec, q, codes = setup(po.L2, seed=6)
    n = vec.shape[0] - 1
    kw = dict(R=12, R_max=16, L_build=32, alpha=1.2, starts=[n], ids=np.arange(n), batch=1)
    a_fp, d_fp = po.build_graph(po.L2, vec, **kw)
    a_s, d_s = po.build_graph(po.L2, vec, sph=(q, codes), **kw)
    assert any(list(a_fp[i, :d_fp[i]]) != list(a_s[i, :d_s[i]]) for i in range(n))
