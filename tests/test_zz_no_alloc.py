"""The allocation-free batched search (dkm_search_batch_device_noalloc) against the
allocating one (dkm_search_batch_device): identical ids, distance bytes, counts, cmps and
hops, on two calls in a row (the second smaller, reusing the same scratch array), for full
precision and Hybrid PQ; launches too big for the scratch array are refused."""
import ctypes as C

import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd.quantization import _Dev
from oracle import pyoracle as po
from test_gpu_search import dataset, log, make_index

pytestmark = pytest.mark.gpu

CAPACITY = 12


def _outs(lib, nq, kst):
    return [_Dev(lib, nbytes=4 * nq * kst), _Dev(lib, nbytes=4 * nq * kst),
            _Dev(lib, nbytes=4 * nq), _Dev(lib, nbytes=4 * nq), _Dev(lib, nbytes=4 * nq)]


def _get(o, nq, kst):
    return (o[0].get(np.uint32, (nq, kst)), o[1].get(np.float32, (nq, kst)),
            o[2].get(np.uint32, nq), o[3].get(np.uint32, nq), o[4].get(np.uint32, nq))


def _run(ix, fn, strategy, q, ks, Ls, max_k, max_L, kst, status=None):
    lib = ix.lib
    nq = q.shape[0]
    dq, dkk, dL = _Dev(lib, q), _Dev(lib, ks), _Dev(lib, Ls)
    o = _outs(lib, nq, kst)
    args = [ix._h, int(strategy), dq.ptr, nq, dkk.ptr, dL.ptr, max_k, max_L, 1, kst,
            o[0].ptr, o[1].ptr, o[2].ptr, o[3].ptr, o[4].ptr]
    if fn == "noalloc":
        st = lib.dkm_search_batch_device_noalloc(*args, status.ptr if status else None)
    else:
        st = lib.dkm_search_batch_device(*args)
    assert st == 0, lib.dkm_last_error()
    assert lib.dkm_index_synchronize(ix._h) == 0
    return _get(o, nq, kst)


def _same(a, b, ks):
    ids_a, d_a, c_a, cm_a, hp_a = a
    ids_b, d_b, c_b, cm_b, hp_b = b
    assert np.array_equal(c_a, c_b) and np.array_equal(cm_a, cm_b) and np.array_equal(hp_a, hp_b)
    for i in range(len(c_a)):
        n = c_a[i]
        assert np.array_equal(ids_a[i, :n], ids_b[i, :n])
        assert d_a[i, :n].tobytes() == d_b[i, :n].tobytes()


@pytest.mark.parametrize("strategy", [dk.FullPrecision, dk.Hybrid])
def test_noalloc_matches_allocating_search_twice(strategy):
    log("noalloc strategy=%d" % int(strategy))
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    if strategy == dk.Hybrid:
        from diskann_amd import quantization as qz
        off = qz.chunk_offsets(vec.shape[1], 32)
        piv = vec[np.random.default_rng(1).choice(vec.shape[0], 256, replace=False)]
        ix.set_pq(piv, off)
        ix.encode()
    rng = np.random.default_rng(5)
    nq = qs.shape[0]
    ks = rng.choice(np.array([1, 10, 20], np.uint32), nq).astype(np.uint32)
    Ls = np.maximum(rng.integers(16, 129, nq), ks).astype(np.uint32)
    ref = _run(ix, "alloc", strategy, qs, ks, Ls, 20, 128, 20)
    status = _Dev(ix.lib, nbytes=4 * nq)
    got = _run(ix, "noalloc", strategy, qs, ks, Ls, 20, 128, 20, status)
    _same(ref, got, ks)
    assert not status.get(np.uint32, nq).any()
    # second call: fewer queries, different order, no caller status buffer
    sub = np.arange(nq)[::-3][:17]
    got2 = _run(ix, "noalloc", strategy, np.ascontiguousarray(qs[sub]), ks[sub].copy(),
                Ls[sub].copy(), 20, 128, 20)
    _same(tuple(a[sub] for a in ref), got2, ks[sub])


def test_noalloc_refuses_launches_beyond_its_scratch():
    log("noalloc capacity")
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    lib = ix.lib
    nq = 2000  # 2000 x 16384 visited slots x 4 bytes > 64 MiB
    q = np.ascontiguousarray(np.resize(qs, (nq, qs.shape[1])))
    ks = np.full(nq, 10, np.uint32)
    Ls = np.full(nq, 50, np.uint32)
    dq, dkk, dL = _Dev(lib, q), _Dev(lib, ks), _Dev(lib, Ls)
    o = _outs(lib, nq, 10)
    st = lib.dkm_search_batch_device_noalloc(ix._h, int(dk.FullPrecision), dq.ptr, nq, dkk.ptr,
                                             dL.ptr, 10, 50, 1, 10, o[0].ptr, o[1].ptr, o[2].ptr,
                                             o[3].ptr, o[4].ptr, None)
    assert st == CAPACITY
    # a smaller visited capacity fits the same launch, and it matches the allocating path
    assert lib.dkm_set_visited_capacity(ix._h, 4096) == 0
    st = lib.dkm_search_batch_device_noalloc(ix._h, int(dk.FullPrecision), dq.ptr, nq, dkk.ptr,
                                             dL.ptr, 10, 50, 1, 10, o[0].ptr, o[1].ptr, o[2].ptr,
                                             o[3].ptr, o[4].ptr, None)
    assert st == 0
    assert lib.dkm_index_synchronize(ix._h) == 0
    got = _get(o, nq, 10)
    ref = _run(ix, "alloc", dk.FullPrecision, q, ks, Ls, 10, 50, 10)
    _same(ref, got, ks)
