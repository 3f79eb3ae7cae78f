"""Large-dimension k-means on the GPU: the FP32 path against the oracle's Lloyd
(lloyds.rs restated: assignment and f64 row-order update, bit-exact), the BF16 MFMA path
against the FP32 path (exactly equal where every product and sum is exact in bf16/f32,
near-equal on real-valued clustered data)."""
import numpy as np
import pytest

from diskann_amd import quantization as qz
from oracle import pyoracle as po
from test_gpu_search import log

pytestmark = pytest.mark.gpu


def clustered(n, dim, k, seed):
    rng = np.random.default_rng(seed)
    cen = rng.standard_normal((k, dim)).astype(np.float32)
    x = cen[rng.integers(0, k, n)] + 0.3 * rng.standard_normal((n, dim)).astype(np.float32)
    return x.astype(np.float32), rng


@pytest.mark.parametrize("n,dim,K", [(3000, 96, 50), (600, 768, 64), (777, 37, 13)])
def test_f32_assign_matches_oracle(n, dim, K):
    log("kmeans f32 assign n=%d dim=%d K=%d" % (n, dim, K))
    x, rng = clustered(n, dim, K, 1)
    c = x[rng.choice(n, K, replace=False)]
    a, _ = qz.kmeans_assign(x, c, qz.F32)
    ao, _ = po.lloyd_assign(x, c)
    assert np.array_equal(a, ao)


@pytest.mark.parametrize("n,dim,K", [(2000, 64, 40), (500, 50, 70), (300, 768, 130),
                                     (1000, 256, 300), (129, 384, 33), (70, 512, 64),
                                     (257, 100, 1), (150, 800, 40)])
def test_bf16_assign_exact_on_integer_data(n, dim, K):
    # small integers: every bf16 input, product and partial sum is exact, so the MFMA
    # scores equal the FP32 ones and the arg-min (ties to the first centre) must agree.
    # Dims up to 768 take the register-resident kernel (padded to 64 .. 768), 800 the
    # LDS-tiled one; K = 1, 33, 300 leave partial centre tiles.
    log("kmeans bf16 exact n=%d dim=%d K=%d" % (n, dim, K))
    rng = np.random.default_rng(2)
    x = rng.integers(-4, 5, (n, dim)).astype(np.float32)
    c = rng.integers(-4, 5, (K, dim)).astype(np.float32)
    a16, d16 = qz.kmeans_assign(x, c, qz.BF16)
    ao, _ = po.lloyd_assign(x, c)
    assert np.array_equal(a16, ao)
    # the reported score is c.c - 2 x.c of the chosen centre (exact here)
    cx = c.astype(np.float64)[a16]
    want = (cx * cx).sum(1) - 2.0 * (x.astype(np.float64) * cx).sum(1)
    assert np.array_equal(d16.astype(np.float64), want)


def test_bf16_assign_close_on_clustered_data():
    x, rng = clustered(4000, 768, 64, 3)
    c = x[rng.choice(4000, 64, replace=False)]
    a16, _ = qz.kmeans_assign(x, c, qz.BF16)
    a32, _ = qz.kmeans_assign(x, c, qz.F32)
    agree = float(np.mean(a16 == a32))
    # where they differ, the bf16 choice is within a tiny relative gap of the optimum (FP64)
    d64 = ((x.astype(np.float64)[:, None, :] - c.astype(np.float64)[None, :, :]) ** 2).sum(-1) \
        if x.shape[0] * c.shape[0] * x.shape[1] < 3e8 else None
    assert agree >= 0.99, agree
    if d64 is not None:
        rows = np.nonzero(a16 != a32)[0]
        best = d64[np.arange(len(x)), a32]
        got = d64[np.arange(len(x)), a16]
        assert np.all(got[rows] <= best[rows] * (1 + 1e-2))


def test_f32_kmeans_iterations_match_oracle():
    log("kmeans f32 iterations")
    x, rng = clustered(2500, 128, 32, 4)
    init = x[rng.choice(2500, 32, replace=False)].copy()
    c_gpu, a_gpu, _, _ = qz.kmeans(x, init, 3, qz.F32)
    c = init.copy()
    for _ in range(3):
        a, _ = po.lloyd_assign(x, c)
        c = po.lloyd_update(x, a, 32)
    assert np.array_equal(a_gpu, a)
    assert c_gpu.tobytes() == c.tobytes()


def test_update_empty_cluster_is_zero():
    x = np.random.default_rng(5).standard_normal((100, 16)).astype(np.float32)
    init = np.vstack([x[:3], np.full((1, 16), 1e6, np.float32)])  # the last centre wins nothing
    c_gpu, a_gpu, _, _ = qz.kmeans(x, init, 1, qz.F32)
    assert not np.any(a_gpu == 3)
    assert not c_gpu[3].any()
    assert c_gpu.tobytes() == po.lloyd_update(x, a_gpu, 4).tobytes()


def test_bf16_kmeans_inertia_close_to_f32():
    x, rng = clustered(5000, 768, 48, 6)
    init = x[rng.choice(5000, 48, replace=False)].copy()
    c32, a32, _, _ = qz.kmeans(x, init, 5, qz.F32)
    c16, a16, _, _ = qz.kmeans(x, init, 5, qz.BF16)

    def inertia(c, a):
        return float(((x.astype(np.float64) - c.astype(np.float64)[a]) ** 2).sum())
    assert inertia(c16, a16) <= inertia(c32, a32) * 1.01
