"""GPU serving loop (dkm_server_*): a request served by the continuous-batching lanes
(batched search launches on their own streams) must return exactly what a batched search
of that query alone returns (ids, f32 distances, cmps, hops), whatever else is in flight,
with mixed k and L per request."""
import time

import numpy as np
import pytest

import diskann_amd as dk
from oracle import pyoracle as po
from test_gpu_search import dataset, log, make_index, pq_setup

pytestmark = pytest.mark.gpu


def drain(sv, want, limit_s=30.0):
    """Poll until `want` results have arrived (bounded wall time)."""
    got = {}
    t0 = time.time()
    while len(got) < want:
        tags, ids, ds, cnt, cmps, hops, lat = sv.poll()
        for i in range(len(tags)):
            got[int(tags[i])] = (ids[i, :cnt[i]].copy(), ds[i, :cnt[i]].copy(), int(cmps[i]),
                                 int(hops[i]), float(lat[i]))
        if time.time() - t0 > limit_s:
            raise AssertionError("served %d of %d requests in %.0f s" % (len(got), want, limit_s))
        if len(got) < want:
            time.sleep(0.0005)
    return got


@pytest.mark.parametrize("strategy", [dk.FullPrecision, dk.Hybrid, dk.Quantized])
def test_served_equals_oracle_mixed_k_L(strategy):
    log("serve strategy=%d" % strategy)
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    piv, off = pq_setup(vec, 32)
    ix.set_pq(piv, off)
    ix.encode()
    codes = ix.get_codes()
    oix = po.OracleIndex(po.L2, vec, adj, deg, [vec.shape[0] - 1], codes, piv, off)
    mode = {dk.FullPrecision: po.MODE_FULL, dk.Hybrid: po.MODE_HYBRID,
            dk.Quantized: po.MODE_PQ_ONLY}[strategy]
    rng = np.random.default_rng(11)
    ks = rng.choice([1, 5, 10], qs.shape[0])
    Ls = np.maximum(ks, rng.integers(8, 128, qs.shape[0]))
    with dk.Server(ix, strategy, max_k=10, max_L=128, beam_width=2, capacity=256,
                   slots=32) as sv:
        for i in range(qs.shape[0]):
            sv.submit(qs[i], int(ks[i]), int(Ls[i]), tag=1000 + i)
        got = drain(sv, qs.shape[0])
        assert sv.pending() == 0
    for i in range(qs.shape[0]):
        ids, ds, cm, hp = oix.search(qs[i], int(ks[i]), int(Ls[i]), 2, mode)
        g = got[1000 + i]
        assert list(g[0]) == list(ids), i
        assert g[1].tobytes() == np.asarray(ds, np.float32).tobytes(), i
        assert (g[2], g[3]) == (cm, hp), i


def test_served_equals_batch_across_idle_gaps():
    # bursts separated by idle time: every burst is launched and served the same way
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    ref = ix.search_batch(qs, 10, 64, 1, dk.FullPrecision)
    with dk.Server(ix, dk.FullPrecision, max_k=10, max_L=64, capacity=16, slots=8) as sv:
        for rnd in range(3):
            for i in range(rnd * 8, rnd * 8 + 8):
                sv.submit(qs[i], 10, 64, tag=i)
            got = drain(sv, 8)
            for i, g in got.items():
                assert list(g[0]) == list(ref.ids[i, :ref.counts[i]])
                assert g[1].tobytes() == ref.distances[i, :ref.counts[i]].tobytes()
                assert (g[2], g[3]) == (ref.cmps[i], ref.hops[i])
                assert g[4] > 0
            time.sleep(0.1)


def test_served_equals_batch_under_continuous_traffic():
    # small bursts with polls in between keep several launches in flight on different
    # streams (slots 16 -> 4 per launch); every result must equal the batched search
    log("serve continuous traffic")
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    ref = ix.search_batch(qs, 10, 64, 1, dk.FullPrecision)
    nreq = 10 * qs.shape[0]
    got = {}
    with dk.Server(ix, dk.FullPrecision, max_k=10, max_L=64, capacity=64, slots=16) as sv:
        i = 0
        t0 = time.time()
        while len(got) < nreq:
            if i < nreq:
                i += sv.submit_batch(qs[[j % qs.shape[0] for j in range(i, min(nreq, i + 8))]],
                                     10, 64, np.arange(i, min(nreq, i + 8)))
            tags, ids, ds, cnt, cmps, hops, lat = sv.poll()
            for j in range(len(tags)):
                got[int(tags[j])] = (ids[j, :cnt[j]].copy(), ds[j, :cnt[j]].copy(), int(cmps[j]),
                                     int(hops[j]))
            assert time.time() - t0 < 60, "served %d of %d" % (len(got), nreq)
    for tag, g in got.items():
        q = tag % qs.shape[0]
        assert list(g[0]) == list(ref.ids[q, :ref.counts[q]]), tag
        assert g[1].tobytes() == ref.distances[q, :ref.counts[q]].tobytes(), tag
        assert (g[2], g[3]) == (ref.cmps[q], ref.hops[q]), tag


def test_server_admission_and_validation():
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    with dk.Server(ix, dk.FullPrecision, max_k=10, max_L=64, capacity=4, slots=4) as sv:
        for i in range(4):
            sv.submit(qs[i], 10, 32, tag=i)
        with pytest.raises(dk.CapacityError):
            sv.submit(qs[4], 10, 32, tag=4)     # ring full until results are polled
        with pytest.raises(dk.KnnSearchError):
            sv.submit(qs[4], 0, 32, tag=4)
        with pytest.raises(dk.KnnSearchError):
            sv.submit(qs[4], 10, 5, tag=4)
        with pytest.raises(dk.ANNError):
            sv.submit(qs[4], 10, 65, tag=4)     # above the server's max_L
        assert sv.pending() == 4
        drain(sv, 4)
        sv.submit(qs[4], 10, 32, tag=4)
        assert len(drain(sv, 1)) == 1
    with pytest.raises(dk.ANNError):
        dk.Server(ix, dk.Hybrid, max_k=10, max_L=64)    # no PQ codes on this index


def test_submit_batch_with_device_sized_slots():
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    ref = ix.search_batch(qs, 10, 48, 1, dk.FullPrecision)
    with dk.Server(ix, dk.FullPrecision, max_k=10, max_L=48, capacity=32, slots=0) as sv:
        assert sv.slots >= 256
        acc = sv.submit_batch(qs, 10, 48, np.arange(qs.shape[0]))
        assert acc == 32                        # ring of 32 entries
        got = drain(sv, 32)
        assert sv.submit_batch(qs[32:], 10, 48, np.arange(32, qs.shape[0])) == 32
        got.update(drain(sv, 32))
    assert sorted(got) == list(range(qs.shape[0]))
    for i, g in got.items():
        assert list(g[0]) == list(ref.ids[i, :ref.counts[i]])
        assert (g[2], g[3]) == (ref.cmps[i], ref.hops[i])
