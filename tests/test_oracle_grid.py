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
