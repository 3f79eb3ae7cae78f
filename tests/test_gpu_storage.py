"""Graph files and vector .bin files against an index on the device: the bytes written
follow the reference's graph layout (diskann-providers/src/storage/bin.rs:310-380 with
simple_neighbor_provider.rs:143-185, 313-356), parsed here independently, and an index
loaded from them searches exactly like the original."""
import struct

import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd import storage as st
from oracle import pyoracle as po
from test_gpu_search import dataset, log, make_index

pytestmark = pytest.mark.gpu


def parse_graph(raw):
    size, maxdeg, start, nstart = struct.unpack_from("<QIIQ", raw, 0)
    assert size == len(raw)
    pos, lists = 24, []
    while pos < size:
        (c,) = struct.unpack_from("<I", raw, pos)
        lists.append(list(struct.unpack_from("<%dI" % c, raw, pos + 4)))
        pos += 4 * (1 + c)
    return maxdeg, start, nstart, lists


def test_save_load_graph_inmem_layout(tmp_path):
    log("graph save/load")
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    n = vec.shape[0] - 1
    p = tmp_path / "graph"
    ix.save_graph(p)
    maxdeg, start, nstart, lists = parse_graph(p.read_bytes())
    assert (maxdeg, start, nstart) == (ix.R_max, n, 1)
    assert len(lists) == n + 1
    for i in range(n + 1):
        assert lists[i] == list(adj[i, :deg[i]]), i
    ix2 = dk.DiskANNIndex(po.L2, vec.shape[1], n, ix.R, ix.R_max)
    ix2.set_vectors(vec[:n])
    ix2.set_start_points(vec[n:])
    assert ix2.load_graph(p) == n
    l2, c2 = ix2.get_graph()
    assert np.array_equal(c2, deg)
    r1 = ix.search_batch(qs, 10, 64, 2, dk.FullPrecision)
    r2 = ix2.search_batch(qs, 10, 64, 2, dk.FullPrecision)
    assert np.array_equal(r1.ids, r2.ids) and r1.distances.tobytes() == r2.distances.tobytes()
    assert np.array_equal(r1.cmps, r2.cmps)


def test_save_graph_disk_layout(tmp_path):
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    n = vec.shape[0] - 1
    actual = 17
    p = tmp_path / "disk_graph"
    ix.save_graph_disk(p, actual)
    maxdeg, start, nstart, lists = parse_graph(p.read_bytes())
    assert (start, nstart) == (actual, 0)
    assert len(lists) == n
    assert maxdeg == max(int(d) for d in deg[:n])
    for i in range(n):
        want = [actual if v == n else int(v) for v in adj[i, :deg[i]]]
        assert lists[i] == want, i


def test_load_graph_rejects_mismatch(tmp_path):
    vec, adj, deg, qs = dataset(po.L2)
    ix = make_index(po.L2, vec, adj, deg)
    p = tmp_path / "g"
    ix.save_graph(p)
    small = dk.DiskANNIndex(po.L2, vec.shape[1], 100, ix.R, ix.R_max)
    with pytest.raises(dk.ANNError):
        small.load_graph(p)
    with pytest.raises(dk.StorageError):
        small.load_graph(tmp_path / "missing")


def test_vectors_from_bin(tmp_path):
    vec, adj, deg, qs = dataset(po.L2)
    n = vec.shape[0] - 1
    p = tmp_path / "base.bin"
    st.write_bin(p, vec[:n])
    ix = make_index(po.L2, vec, adj, deg)
    ix2 = dk.DiskANNIndex(po.L2, vec.shape[1], n, ix.R, ix.R_max)
    ix2.set_vectors_from_bin(p)
    ix2.set_start_points(vec[n:])
    ix2.set_graph(adj, deg)
    r1 = ix.search_batch(qs, 10, 32, 1, dk.FullPrecision)
    r2 = ix2.search_batch(qs, 10, 32, 1, dk.FullPrecision)
    assert np.array_equal(r1.ids, r2.ids) and r1.distances.tobytes() == r2.distances.tobytes()
    bad = tmp_path / "bad.bin"
    st.write_bin(bad, vec[:10, :5])
    with pytest.raises(dk.DimensionMismatchError):
        ix2.set_vectors_from_bin(bad)
