"""Device-buffer search paths report per-query failures (dkm_search_status,
dkm_search_graph_status): k / L read on the device are checked against Knn::new's rules
(diskann/src/graph/search/knn_search.rs:101-106) and the launch's max_L / k_stride before a
query runs (a refused query writes nothing but its zero count, so neighbouring outputs stay
intact), and a visited-set overflow or iteration bound comes back as an error instead of a
silently truncated result."""
import ctypes as C

import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd.errors import ANNError, CapacityError
from diskann_amd.quantization import _Dev
from oracle import pyoracle as po
from test_gpu_search import dataset, log, make_index

pytestmark = pytest.mark.gpu

INVALID_ARG, CAPACITY = 1, 12


def _launch(ix, qs, ks, Ls, max_k, max_L, k_stride, strategy=dk.FullPrecision):
    lib = ix.lib
    nq = qs.shape[0]
    dq, dk_, dL = _Dev(lib, qs), _Dev(lib, ks.astype(np.uint32)), _Dev(lib, Ls.astype(np.uint32))
    sentinel = np.full((nq, k_stride), 0xABCDABCD, np.uint32)
    dids = _Dev(lib, sentinel)
    dd = _Dev(lib, sentinel.view(np.float32).copy())
    dcnt = _Dev(lib, np.full(nq, 77, np.uint32))
    dcm = _Dev(lib, nbytes=4 * nq)
    dhp = _Dev(lib, nbytes=4 * nq)
    st = lib.dkm_search_batch_device(ix._h, int(strategy), dq.ptr, nq, dk_.ptr, dL.ptr, max_k,
                                     max_L, 1, k_stride, dids.ptr, dd.ptr, dcnt.ptr, dcm.ptr,
                                     dhp.ptr)
    assert st == 0
    first = C.c_int64(-2)
    st = lib.dkm_search_status(ix._h, C.byref(first))
    return (st, first.value, dids.get(np.uint32, (nq, k_stride)),
            dd.get(np.float32, (nq, k_stride)), dcnt.get(np.uint32, nq),
            dcm.get(np.uint32, nq), dhp.get(np.uint32, nq))


def test_device_path_refuses_bad_k_and_L_without_touching_neighbours():
    log("device status: bad k / L")
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    oix = po.OracleIndex(po.L2, vec, adj, deg, [vec.shape[0] - 1])
    nq = 12
    q = qs[:nq]
    ks = np.full(nq, 10, np.int64)
    Ls = np.full(nq, 40, np.int64)
    max_k, max_L, k_stride = 10, 64, 10
    bad = {3: (10, 65),          # L above the launch's max_L (LDS queue capacity)
           5: (11, 40),          # k above the output stride
           7: (0, 40),           # k = 0
           8: (10, 0),           # L = 0
           10: (20, 12),         # L < k
           11: (10, 1 << 31)}    # absurd L
    for i, (kk, ll) in bad.items():
        ks[i], Ls[i] = kk, ll
    st, first, ids, ds, cnt, cm, hp = _launch(ix, q, ks, Ls, max_k, max_L, k_stride)
    assert st == INVALID_ARG and first == 3
    for i in range(nq):
        if i in bad:
            assert cnt[i] == 0 and cm[i] == 0 and hp[i] == 0
            assert (ids[i] == 0xABCDABCD).all()      # nothing written to its output row
            continue
        oids, ods, ocm, ohp = oix.search(q[i], 10, 40, 1)
        assert list(ids[i, :cnt[i]]) == list(oids)
        assert ds[i, :cnt[i]].tobytes() == ods.tobytes()
        assert (cm[i], hp[i]) == (ocm, ohp)
    # a clean launch afterwards reports success
    st, first, ids, ds, cnt, cm, hp = _launch(ix, q, np.full(nq, 10), np.full(nq, 40), max_k,
                                              max_L, k_stride)
    assert st == 0 and first == -1


def test_visited_overflow_is_reported_on_device_and_graph_paths():
    log("device status: visited overflow")
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    ix.set_tuning(visited_lds_slots=256)
    lib = ix.lib
    assert lib.dkm_set_visited_capacity(ix._h, 256) == 0
    nq = 8
    ks, Ls = np.full(nq, 10), np.full(nq, 300)
    st, first, *_ = _launch(ix, qs[:nq], ks, Ls, 10, 300, 10)
    assert st == CAPACITY and first == 0
    with pytest.raises(CapacityError):
        ix.search_batch(qs[:nq], 10, 300)
    with pytest.raises(CapacityError):
        dk.SearchGraph(ix, qs[:nq], 10, 300).run()
    # automatic capacity again: the same searches succeed and match the oracle
    assert lib.dkm_set_visited_capacity(ix._h, 0) == 0
    st, first, ids, ds, cnt, cm, hp = _launch(ix, qs[:nq], ks, Ls, 10, 300, 10)
    assert st == 0
    oix = po.OracleIndex(po.L2, vec, adj, deg, [vec.shape[0] - 1])
    for i in range(nq):
        oids, ods, ocm, ohp = oix.search(qs[i], 10, 300, 1)
        assert list(ids[i, :cnt[i]]) == list(oids) and (cm[i], hp[i]) == (ocm, ohp)
    g = dk.SearchGraph(ix, qs[:nq], 10, 300)
    r = g.run()
    assert list(r.ids[0, :r.counts[0]]) == list(oix.search(qs[0], 10, 300, 1)[0])
    g.close()


def test_bad_arguments_of_the_capacity_knob():
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    for bad in (-1, 100, 1000, 1 << 25):
        assert ix.lib.dkm_set_visited_capacity(ix._h, bad) == INVALID_ARG
    with pytest.raises(ANNError):
        from diskann_amd.errors import check
        check(ix.lib.dkm_set_visited_capacity(ix._h, 1000), ix.lib)
