"""Pin the oracle's Vamana build (insert / multi_insert / RobustPrune) against the
reference's golden grid_insert cases: after inserting every grid point (one at a time,
or in batches with intra-batch candidates none / max(4) / all), the two golden searches
must return the same ids, distances, comparisons and hops
(diskann/src/graph/test/cases/grid_insert.rs, generated JSON under
diskann/test/generated/graph/test/cases/grid_insert/).

Configuration restated from grid_insert.rs: max degree 2*dim, pruned degree
max(2, 2*dim - 2), L_build 100, alpha 1.2, L2, start point [size]*dim with no edges."""
import json
import os

import numpy as np
import pytest

from gridutil import grid
from oracle import pyoracle as po

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "grid_cases.json")))


def case_params(name):
    head, ibc = name.split("/")
    parts = head.split("_")
    dim, size = int(parts[1]), int(parts[2])
    batch = 1 if parts[3] == "single" else int(parts[4])
    intra = {"ibc_none": 0, "ibc_all": -1}.get(ibc)
    if intra is None:
        intra = int(ibc.split("_")[-1])
    return dim, size, batch, intra


def build_grid_index(dim, size, batch, intra):
    vec, _, _, starts = grid(dim, size)
    n = vec.shape[0] - 1
    R_max = 2 * dim
    R = min(max(2, R_max - 2), R_max)
    adj, deg = po.build_graph(po.L2, vec, R=R, R_max=R_max, L_build=100, alpha=1.2,
                              starts=starts, ids=np.arange(n), batch=batch, intra_batch=intra,
                              max_backedges=R)
    return vec, adj, deg, starts


@pytest.mark.parametrize("name", sorted(FIX["grid_insert"]))
def test_grid_insert_golden(name):
    """Distances of the top-10 after the build must equal the reference's in every case.
    Ids, comparisons and hops must match exactly on the 1-D grids, whose candidate pools
    have no distance ties. On the 3-D/4-D grids the reference orders equal-distance
    candidates with Rust's select_nth_unstable / sort_unstable
    (diskann/src/graph/internal/sorted_neighbors.rs:26-44), an order the language leaves
    unspecified; this restatement sorts stably, so ids there may be a permutation within
    ties (checked: same multiset of distances)."""
    c = FIX["grid_insert"][name]
    dim, size, batch, intra = case_params(name)
    vec, adj, deg, starts = build_grid_index(dim, size, batch, intra)
    ix = po.OracleIndex(po.L2, vec, adj, deg, starts)
    for s in c["searches"]:
        ids, ds, cm, hp = ix.search(np.array(s["query"], np.float32), 10, 10, 1)
        assert [float(d) for d in ds] == [float(e[1]) for e in s["results"]], name
        if dim == 1:
            assert [int(i) for i in ids] == [int(e[0]) for e in s["results"]], name
            assert (cm, hp) == (s["comparisons"], s["hops"]), name


def test_prune_list_on_a_line_keeps_the_nearest_on_each_side():
    # location 0 at the origin, candidates on one line: for any candidate beyond the
    # nearest one on its side, that nearest one occludes it by a factor d_ik / d_jk > alpha
    # (occlude_list, diskann/src/graph/index.rs:2570-2782), so only the two nearest remain;
    # saturation then appends the rest in pool (distance) order
    xs = np.array([0.0, 3.0, -2.0, 1.0, 5.0, -1.0, 2.0, -4.0], np.float32)
    vec = np.zeros((len(xs), 4), np.float32)
    vec[:, 0] = xs
    lst = np.arange(1, len(xs), dtype=np.uint32)
    got = po.prune_list(po.L2, vec, 0, lst, 6, 1.2)
    assert list(got) == [3, 5]                        # x = 1, then x = -1 (list order on ties)
    sat = po.prune_list(po.L2, vec, 0, lst, 6, 1.2, saturate=True)
    assert list(sat) == [3, 5, 2, 6, 1, 7]            # then x = -2, 2, 3, -4 (distance, list order)
    assert list(po.prune_list(po.L2, vec, 0, np.array([0], np.uint32), 6, 1.2)) == []
    assert list(po.prune_list(po.L2, vec, 0, np.array([], np.uint32), 6, 1.2)) == []


def test_prune_list_equals_occlude_list_on_the_sorted_pool():
    rng = np.random.default_rng(4)
    vec = rng.standard_normal((300, 16)).astype(np.float32)
    lst = rng.choice(np.arange(1, 300), 120, replace=False).astype(np.uint32)
    d = np.array([po.distance(po.L2, vec[0], vec[i]) for i in lst], np.float32)
    order = np.argsort(d, kind="stable")
    want = po.occlude_list(po.L2, vec, lst[order], d[order], 0, 12, 1.2)
    assert np.array_equal(po.prune_list(po.L2, vec, 0, lst, 12, 1.2), want)
