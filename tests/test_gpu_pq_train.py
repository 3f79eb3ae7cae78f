"""GPU PQ training (dkm_pq_train / dkm_pq_train_draws: chunk offsets, k-means++ seeding,
Lloyd assign and update) against the oracle's train driver
(diskann-quantization/src/product/train.rs:109-231, plusplus.rs:238-498,
lloyds.rs:28-426) given the same k-means++ draws: pivots byte for byte, including the
training sets of the served 10M x 96 (PQ 48) and batched 1M x 128 (PQ 32) configurations
(100k-row samples, 10 Lloyd iterations), uneven chunk widths, K < 256, chunks wider than
the Lloyd update's register path, and the k-means++ failure cases."""
import numpy as np
import pytest

from diskann_amd import quantization as qz
from diskann_amd.errors import PQError
from oracle import pyoracle as po
from test_gpu_search import log

pytestmark = pytest.mark.gpu


def gmm(n, dim, comps, sigma, seed):
    rng = np.random.default_rng(seed)
    cen = rng.standard_normal((comps, dim)).astype(np.float32)
    x = cen[rng.integers(0, comps, n)] + (sigma * rng.standard_normal((n, dim))).astype(np.float32)
    return x.astype(np.float32), rng


@pytest.mark.parametrize("n,dim,M,K,iters", [
    (100_000, 96, 48, 256, 10),     # served search configuration: PQ 48 x 256, dsub 2
    (100_000, 128, 32, 256, 10),    # batched PQ configuration: PQ 32 x 256, dsub 4
    (5000, 50, 7, 256, 5),          # uneven chunks (8 x 1 extra ... widths 7 and 8)
    (3000, 96, 12, 64, 3),          # dsub 8, K < 256
    (2000, 200, 4, 100, 2),         # dsub 50: generic assignment path
    (1500, 1536, 48, 256, 2),       # dsub 32 (config-5 dimension)
    (1500, 128, 2, 256, 2),         # dsub 64: the widest chunk (Lloyd update row tile 192)
])
def test_pq_train_matches_oracle(n, dim, M, K, iters):
    log("pq train n=%d dim=%d M=%d K=%d iters=%d" % (n, dim, M, K, iters))
    x, rng = gmm(n, dim, 256, 0.5, n + dim)
    u = rng.random((M, K))
    piv, off = qz.train_pq_draws(x, M, u, num_centers=K, lloyd_iters=iters)
    assert list(off) == list(po.chunk_offsets(dim, M))
    want = po.pq_train(x, off, K, iters, u)
    assert piv.tobytes() == want.tobytes()


def test_seeded_training_uses_the_documented_stream():
    log("pq train seeded stream")
    x, _ = gmm(4000, 32, 64, 0.5, 5)
    piv, off = qz.train_pq(x, 8, 256, 3, seed=0xD15C3004)
    u = qz.train_draws(0xD15C3004, 8, 256)
    assert piv.tobytes() == po.pq_train(x, off, 256, 3, u).tobytes()


def test_kmeanspp_only_matches_oracle_seeding():
    # zero Lloyd iterations: the pivots are the k-means++ centres themselves
    log("pq train k-means++ only")
    x, rng = gmm(20_000, 64, 128, 0.3, 9)
    u = rng.random((16, 256))
    piv, off = qz.train_pq_draws(x, 16, u, num_centers=256, lloyd_iters=0)
    for m in range(16):
        xc = np.ascontiguousarray(x[:, off[m]:off[m + 1]])
        c, sel, reason = po.kmeanspp(xc, 256, u[m])
        assert reason == po.KPP_OK and sel == 256
        assert piv[:, off[m]:off[m + 1]].tobytes() == c.tobytes()


def test_recoverable_failures_leave_zero_centres():
    log("pq train recoverable failures")
    rng = np.random.default_rng(3)
    # fewer rows than centres: DatasetTooSmall, every row picked, the rest zero
    x = rng.standard_normal((100, 16)).astype(np.float32)
    u = rng.random((4, 256))
    piv, off = qz.train_pq_draws(x, 4, u, num_centers=256, lloyd_iters=2)
    assert piv.tobytes() == po.pq_train(x, off, 256, 2, u).tobytes()
    # duplicated rows: InsufficientDiversity once the distinct rows are used up
    y = np.repeat(rng.standard_normal((10, 8)).astype(np.float32), 30, axis=0)
    u = rng.random((2, 64))
    piv, off = qz.train_pq_draws(y, 2, u, num_centers=64, lloyd_iters=2)
    assert piv.tobytes() == po.pq_train(y, off, 64, 2, u).tobytes()


def test_infinite_total_is_an_error():
    log("pq train SawInfinity")
    x = np.zeros((500, 8), np.float32)
    x[:, :] = np.arange(500, dtype=np.float32)[:, None]
    x[123, :4] = 3e38
    u = np.zeros((2, 16))
    with pytest.raises(PQError):
        qz.train_pq_draws(x, 2, u, num_centers=16, lloyd_iters=1)
