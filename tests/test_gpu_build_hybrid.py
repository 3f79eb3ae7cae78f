"""GPU Hybrid prune (HybridPruneAccessor, diskann-providers/src/model/graph/provider/async_/
inmem/product.rs:143-230; HybridComputer, async_/distances.rs:118-165) against the
oracle's Hybrid build on the same rows, codes and pivots: identical adjacency lists.

Cases: max_fp 0 (every id quantized, extras and back edges code vs code), 2 (self and one
more id full precision), 32, and Hybrid::new(None) (every filled id full precision, the
record beyond max_occlusion quantized); single inserts, batches with all intra-batch
candidates, and batches with Max(8) intra-batch candidates (the bootstrap prune runs).
Plus: rows that equal the vectors their codes reconstruct give the full-precision graph
for every max_fp, and the error when codes are missing."""
import numpy as np
import pytest

import diskann_amd as dk
from oracle import pyoracle as po
from test_gpu_build import data

pytestmark = pytest.mark.gpu

N, R, R_MAX, L = 600, 16, 20, 32


def pq_tables(vec, dim, M, seed=5):
    rng = np.random.default_rng(seed)
    piv = (vec[rng.integers(0, vec.shape[0], 256)]
           + 0.05 * rng.standard_normal((256, dim))).astype(np.float32)
    off = po.chunk_offsets(dim, M)
    codes = po.pq_encode(piv, off, vec)
    return piv, off, codes


def gpu_hybrid_build(metric, vec, piv, off, codes, max_fp, batch, intra, max_occlusion=750):
    n = vec.shape[0] - 1
    ix = dk.DiskANNIndex(metric, vec.shape[1], n, R, R_MAX, num_start=1)
    ix.set_vectors(vec[:n])
    ix.set_start_points(vec[n:])
    ix.set_pq(piv, off)
    ix.set_codes(codes)
    if max_fp is not False:
        ix.set_build_hybrid(True, max_fp)
    ix.build(np.arange(n), batch=batch, intra_batch=intra, L_build=L, alpha=1.2,
             max_occlusion=max_occlusion)
    return ix.get_graph()


def oracle_hybrid_build(metric, vec, piv, off, codes, max_fp, batch, intra, max_occlusion=750):
    n = vec.shape[0] - 1
    m = None if max_fp is False else (2**31 - 1 if max_fp is None else max_fp)
    return po.build_graph(metric, vec, R=R, R_max=R_MAX, L_build=L, alpha=1.2, starts=[n],
                          ids=np.arange(n), batch=batch, intra_batch=intra,
                          max_occlusion=max_occlusion, max_fp=m, codes=codes, pivots=piv,
                          pq_offsets=off)


def same_graph(lists, counts, adj, deg):
    assert np.array_equal(counts, deg)
    for i in range(adj.shape[0]):
        assert list(lists[i, :counts[i]]) == list(adj[i, :deg[i]]), i


@pytest.mark.parametrize("metric,dim,M", [(po.L2, 32, 8), (po.INNER_PRODUCT, 24, 6),
                                          (po.COSINE, 20, 5), (po.L2, 30, 7)])
@pytest.mark.parametrize("max_fp", [0, 2, 32, None])
@pytest.mark.parametrize("batch,intra", [(1, 0), (64, -1), (64, 8)])
def test_gpu_hybrid_build_matches_oracle(metric, dim, M, max_fp, batch, intra):
    print("hybrid build metric=%d dim=%d M=%d max_fp=%s batch=%d intra=%d"
          % (metric, dim, M, max_fp, batch, intra), flush=True)
    vec = data(metric, N, dim, seed=2)
    piv, off, codes = pq_tables(vec, dim, M)
    adj, deg = oracle_hybrid_build(metric, vec, piv, off, codes, max_fp, batch, intra)
    lists, counts = gpu_hybrid_build(metric, vec, piv, off, codes, max_fp, batch, intra)
    same_graph(lists, counts, adj, deg)


def test_gpu_hybrid_small_max_occlusion():
    """max_occlusion below the record length: Hybrid::new(None) quantizes the record
    entries the fill's take(max_occlusion) leaves out."""
    print("hybrid build max_occlusion=24", flush=True)
    vec = data(po.L2, N, 32, seed=4)
    piv, off, codes = pq_tables(vec, 32, 8)
    adj, deg = oracle_hybrid_build(po.L2, vec, piv, off, codes, None, 32, -1, max_occlusion=24)
    lists, counts = gpu_hybrid_build(po.L2, vec, piv, off, codes, None, 32, -1, max_occlusion=24)
    same_graph(lists, counts, adj, deg)
    fa, fd = oracle_hybrid_build(po.L2, vec, piv, off, codes, False, 32, -1, max_occlusion=24)
    assert not (np.array_equal(fd, deg) and np.array_equal(fa, adj))


@pytest.mark.parametrize("max_fp", [0, 2])
def test_gpu_hybrid_exact_codes_give_fp_graph(max_fp):
    """Rows equal to the vectors their codes reconstruct on integer pivots: every PQ form
    equals the full-precision distance, so the Hybrid graph is the full-precision one."""
    print("hybrid exact codes max_fp=%d" % max_fp, flush=True)
    dim, M, K = 12, 4, 20
    rng = np.random.default_rng(9)
    piv = rng.integers(-4, 5, (256, dim)).astype(np.float32)  # codes use the first K
    off = po.chunk_offsets(dim, M)
    codes = rng.integers(0, K, (N + 1, M)).astype(np.uint8)
    vec = np.zeros((N + 1, dim), np.float32)
    for c in range(M):
        vec[:, off[c]:off[c + 1]] = piv[codes[:, c], off[c]:off[c + 1]]
    fl, fc = gpu_hybrid_build(po.L2, vec, piv, off, codes, False, 16, -1)
    hl, hc = gpu_hybrid_build(po.L2, vec, piv, off, codes, max_fp, 16, -1)
    same_graph(hl, hc, fl, fc)


def test_gpu_hybrid_needs_codes():
    vec = data(po.L2, 64, 16)
    ix = dk.DiskANNIndex(po.L2, 16, 64, R, R_MAX, num_start=1)
    ix.set_vectors(vec[:64])
    ix.set_start_points(vec[64:])
    ix.set_build_hybrid(True, 4)
    with pytest.raises(dk.ANNError):
        ix.build(np.arange(64), batch=8, L_build=L)
