"""Direct PQ distances on the GPU (dkm_pq_distance_batch) against the oracle's restatement
of FixedChunkPQTable's resumable distances (fixed_chunk_pq_table.rs:32-56, :290-365),
bit for bit: full-precision vs code and code vs code, every metric, even and uneven chunk
partitions (chunks of 1-7, 8-31 and 32+ dimensions reach the remainder, epilogue and main
loop steps), K < 256; plus the seeded-pivot KAT (dynamic.rs:411-474) on the device."""
import numpy as np
import pytest

from diskann_amd import quantization as qz
from diskann_amd.errors import ANNError
from oracle import pyoracle as po
from test_gpu_search import log
from test_oracle_pq_direct import expected_vector, seed_pivot_table

pytestmark = pytest.mark.gpu

METRICS = [po.L2, po.INNER_PRODUCT, po.COSINE, po.COSINE_NORMALIZED]


@pytest.mark.parametrize("dim,M,K", [(128, 32, 256), (50, 19, 200), (128, 5, 256),
                                     (768, 12, 256), (96, 48, 256), (200, 7, 64), (37, 37, 3)])
def test_direct_matches_oracle(dim, M, K):
    log("pq direct dim=%d M=%d K=%d" % (dim, M, K))
    rng = np.random.default_rng(dim * 7 + M)
    piv = rng.standard_normal((K, dim)).astype(np.float32)
    off = po.chunk_offsets(dim, M)
    n, nq, m = 300, 6, 40
    codes = rng.integers(0, K, (n, M)).astype(np.uint8)
    q = rng.standard_normal((nq, dim)).astype(np.float32)
    qc = rng.integers(0, K, (nq, M)).astype(np.uint8)
    ids = rng.integers(0, n, (nq, m)).astype(np.uint32)
    for metric in METRICS:
        got = qz.pq_distance(metric, piv, off, codes, ids, queries=q)
        gqq = qz.pq_distance(metric, piv, off, codes, ids, query_codes=qc)
        for i in range(nq):
            for j in range(m):
                want = np.float32(po.pq_direct(metric, piv, off, q[i], codes[ids[i, j]]))
                assert got[i, j].tobytes() == want.tobytes(), (metric, i, j)
                want = np.float32(po.pq_qq(metric, piv, off, qc[i], codes[ids[i, j]]))
                assert gqq[i, j].tobytes() == want.tobytes(), (metric, i, j)


@pytest.mark.parametrize("dim,pq_chunks", [(20, 7), (200, 7), (20, 20)])
def test_quant_quant_kat_on_device(dim, pq_chunks):
    piv, off = seed_pivot_table(dim, pq_chunks, 20)
    rng = np.random.default_rng(0xd49347d0)
    codes = rng.integers(0, 20, (64, pq_chunks)).astype(np.uint8)
    ids = np.arange(1, 64, dtype=np.uint32)[None, :]
    for metric in (po.L2, po.INNER_PRODUCT, po.COSINE):
        got = qz.pq_distance(metric, piv, off, codes, ids, query_codes=codes[:1])
        v0 = expected_vector(codes[0], off)
        for j in range(ids.shape[1]):
            v1 = expected_vector(codes[ids[0, j]], off)
            assert np.float32(got[0, j]) == np.float32(po.distance(metric, v0, v1))


def test_errors():
    piv = np.zeros((4, 8), np.float32)
    off = po.chunk_offsets(8, 2)
    codes = np.zeros((3, 2), np.uint8)
    with pytest.raises(ANNError):
        qz.pq_distance(po.L2, piv, off, codes, np.array([[3]], np.uint32),
                       queries=np.zeros((1, 8), np.float32))
    bad = codes.copy()
    bad[1, 1] = 4
    with pytest.raises(ANNError):
        qz.pq_distance(po.L2, piv, off, bad, np.array([[0]], np.uint32),
                       queries=np.zeros((1, 8), np.float32))
    with pytest.raises(ANNError):
        qz.pq_distance(po.L2, piv, off, codes, np.array([[0]], np.uint32))
