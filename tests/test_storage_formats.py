"""On-disk formats (host side, no GPU): .bin matrices and PQ pivots files written and read
through the C ABI, checked against an independent reader/writer of the layouts the
reference defines (diskann-utils/src/io.rs:60-100; diskann-providers/src/storage/
pq_storage.rs:93-149,278-349). The reference's own data files under test_data are Git LFS
pointers in this checkout, so the layouts are pinned by the format definitions."""
import struct

import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd import storage as st


def parse_bin(raw, off, dtype):
    r, c = struct.unpack_from("<II", raw, off)
    n = r * c * np.dtype(dtype).itemsize
    return np.frombuffer(raw[off + 8: off + 8 + n], dtype=dtype).reshape(r, c), off + 8 + n


@pytest.mark.parametrize("dtype", [np.float32, np.uint8, np.uint32, np.int8])
def test_bin_roundtrip_and_layout(tmp_path, dtype):
    rng = np.random.default_rng(1)
    a = (rng.standard_normal((37, 13)) * 50).astype(dtype)
    p = tmp_path / "x.bin"
    st.write_bin(p, a)
    raw = p.read_bytes()
    assert len(raw) == 8 + a.nbytes
    assert struct.unpack_from("<II", raw, 0) == (37, 13)
    b, end = parse_bin(raw, 0, dtype)
    assert end == len(raw) and np.array_equal(a, b)
    assert st.read_bin_shape(p) == (37, 13)
    assert np.array_equal(st.read_bin(p, dtype), a)


def test_bin_written_elsewhere_reads_back(tmp_path):
    a = np.arange(24, dtype=np.float32).reshape(6, 4)
    p = tmp_path / "y.bin"
    p.write_bytes(struct.pack("<II", 6, 4) + a.tobytes())
    assert np.array_equal(st.read_bin(p), a)


def test_bin_errors(tmp_path):
    with pytest.raises(dk.StorageError):
        st.read_bin(tmp_path / "missing.bin")
    p = tmp_path / "short.bin"
    p.write_bytes(struct.pack("<II", 10, 10) + b"\0" * 12)
    with pytest.raises(dk.StorageError):
        st.read_bin(p)
    p.write_bytes(b"\1\0")
    with pytest.raises(dk.StorageError):
        st.read_bin_shape(p)


def test_pq_pivots_layout(tmp_path):
    rng = np.random.default_rng(2)
    K, D, M = 256, 40, 7
    piv = rng.standard_normal((K, D)).astype(np.float32)
    off = np.array([0, 6, 12, 18, 24, 30, 35, 40], np.int64)
    p = tmp_path / "pq_pivots.bin"
    st.write_pq_pivots(p, piv, off)
    raw = p.read_bytes()
    tab, _ = parse_bin(raw, 0, np.uint64)
    assert tab.shape == (4, 1)
    o = [int(v) for v in tab[:, 0]]
    assert o[0] == 4096
    assert o[1] == 4096 + 8 + K * D * 4
    assert o[2] == o[1] + 8 + D * 4
    assert o[3] == o[2] + 8 + (M + 1) * 4 == len(raw)
    assert raw[40:4096] == b"\0" * (4096 - 40)
    pv, _ = parse_bin(raw, o[0], np.float32)
    cen, _ = parse_bin(raw, o[1], np.float32)
    co, _ = parse_bin(raw, o[2], np.uint32)
    assert pv.shape == (K, D) and np.array_equal(pv, piv)
    assert cen.shape == (D, 1) and not cen.any()
    assert co.shape == (M + 1, 1) and list(co[:, 0]) == list(off)
    piv2, off2 = st.read_pq_pivots(p)
    assert piv2.tobytes() == piv.tobytes() and list(off2) == list(off)


def test_pq_pivots_centroid_is_folded_on_read(tmp_path):
    # pq_storage.rs:342-346: a non-zero centroid is added to every pivot row on load
    rng = np.random.default_rng(3)
    piv = rng.standard_normal((16, 8)).astype(np.float32)
    cen = rng.standard_normal(8).astype(np.float32)
    p = tmp_path / "c.bin"
    st.write_pq_pivots(p, piv, [0, 4, 8], centroid=cen)
    piv2, off2 = st.read_pq_pivots(p)
    assert piv2.tobytes() == (piv + cen[None, :]).astype(np.float32).tobytes()
    assert list(off2) == [0, 4, 8]


def test_pq_pivots_written_elsewhere_reads_back(tmp_path):
    # a file assembled by hand from the layout definition
    rng = np.random.default_rng(4)
    K, D = 4, 6
    piv = rng.standard_normal((K, D)).astype(np.float32)
    cen = np.zeros(D, np.float32)
    offs = np.array([0, 2, 4, 6], np.uint32)
    blocks = [struct.pack("<II", K, D) + piv.tobytes(), struct.pack("<II", D, 1) + cen.tobytes(),
              struct.pack("<II", 4, 1) + offs.tobytes()]
    c0 = 4096
    c1 = c0 + len(blocks[0])
    c2 = c1 + len(blocks[1])
    c3 = c2 + len(blocks[2])
    head = struct.pack("<II", 4, 1) + struct.pack("<4Q", c0, c1, c2, c3)
    p = tmp_path / "h.bin"
    p.write_bytes(head + b"\0" * (4096 - len(head)) + b"".join(blocks))
    piv2, off2 = st.read_pq_pivots(p)
    assert piv2.tobytes() == piv.tobytes() and list(off2) == [0, 2, 4, 6]


def test_pq_pivots_errors(tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(struct.pack("<II", 3, 1) + b"\0" * 24)
    with pytest.raises(dk.PQError):
        st.read_pq_pivots(p)
    with pytest.raises(dk.StorageError):
        st.read_pq_pivots(tmp_path / "none.bin")


def test_pq_codes_roundtrip(tmp_path):
    codes = np.random.default_rng(5).integers(0, 256, (100, 32), dtype=np.uint8)
    p = tmp_path / "codes.bin"
    st.write_pq_codes(p, codes)
    assert np.array_equal(st.read_pq_codes(p), codes)
