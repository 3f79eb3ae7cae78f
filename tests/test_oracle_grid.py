"""Pin the oracle's queue + beam search against the reference's golden grid searches
(diskann/test/generated/graph/test/cases/grid_search/*.json): ids, f32 distances
(including tie order), comparisons and hops for beam widths 1, 2 and 4."""
import json
import os

import numpy as np
import pytest

from gridutil import grid
from oracle import pyoracle as po

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "grid_cases.json")))


@pytest.mark.parametrize("case", sorted(FIX["grid_search"]))
def test_grid_search_golden(case):
    c = FIX["grid_search"][case]
    vectors, adj, deg, starts = grid(c["grid_dims"], c["grid_size"])
    ix = po.OracleIndex(po.L2, vectors, adj, deg, starts)
    for s in c["searches"]:
        ids, ds, cmps, hops = ix.search(np.array(s["query"], np.float32), c["k"], c["L"],
                                        s["beam_width"])
        exp = s["results"]
        assert len(ids) == s["num_results"]
        assert [int(i) for i in ids] == [int(e[0]) for e in exp]
        assert [float(d) for d in ds] == [float(e[1]) for e in exp]
        assert cmps == s["comparisons"]
        assert hops == s["hops"]


def test_mixed_batch_matches_single_searches():
    """A batch with one k and one L per query (the served request mix) gives each query
    the ids, cmps and hops of its own single search; k > L is refused as Knn::new does
    (diskann/src/graph/search/knn_search.rs:101-106)."""
    rng = np.random.default_rng(1)
    n, d = 3000, 16
    x = rng.standard_normal((n, d)).astype(np.float32)
    med, _ = po.medoid(x)
    vec = np.vstack([x, x[med:med + 1]])
    adj, deg = po.build_graph(po.L2, vec, R=12, R_max=15, L_build=30, alpha=1.2, starts=[n],
                              ids=np.arange(n), batch=64)
    oix = po.OracleIndex(po.L2, vec, adj, deg, [n])
    qs = rng.standard_normal((50, d)).astype(np.float32)
    ks = rng.choice([3, 10], 50).astype(np.uint32)
    Ls = rng.integers(10, 40, 50).astype(np.uint32)
    ids, ds, cm, hp, _ = oix.search_batch_mixed(qs, ks, Ls, 1, po.MODE_FULL, 4)
    for i in range(50):
        a, b, c, e, _, _ = oix.search_batch(qs[i:i + 1], int(ks[i]), int(Ls[i]), 1,
                                            po.MODE_FULL, 1)
        assert np.array_equal(ids[i, :ks[i]], a[0]) and ds[i, :ks[i]].tobytes() == b[0].tobytes()
        assert (cm[i], hp[i]) == (c[0], e[0])
        assert (ids[i, ks[i]:] == 0xFFFFFFFF).all()
    Ls[7] = 2
    with pytest.raises(ValueError):
        oix.search_batch_mixed(qs, ks, Ls, 1, po.MODE_FULL, 4)
