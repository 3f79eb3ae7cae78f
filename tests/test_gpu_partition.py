"""GPU closest centres of the partitioned build (dkm_closest_centers) against the
oracle's compute_closest_centers restatement (diskann-disk/src/utils/math_util.rs:133-385):
identical centre ids for every row, k = 1 and k > 1, dimensions across the kernel's
128-dimension chunks, centre counts across its 32-centre blocks, and exact integer data."""
import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd import quantization as qz
from oracle import pyoracle as po
from test_gpu_search import log

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,dim,K,k", [(5000, 128, 40, 1), (5000, 128, 40, 2), (3000, 96, 64, 3),
                                       (2000, 300, 33, 1), (2000, 768, 70, 4), (777, 5, 7, 7),
                                       (1000, 1536, 16, 16)])
def test_closest_centers_match_oracle(n, dim, K, k):
    log("closest centres n=%d dim=%d K=%d k=%d" % (n, dim, K, k))
    rng = np.random.default_rng(n + dim + K)
    c = rng.standard_normal((K, dim)).astype(np.float32)
    x = (c[rng.integers(0, K, n)] + 0.7 * rng.standard_normal((n, dim))).astype(np.float32)
    got = qz.closest_centers(x, c, k)
    want = po.closest_centers(x, c, k)
    assert np.array_equal(got, want)


def test_closest_centers_integer_data_is_brute_force():
    rng = np.random.default_rng(3)
    x = rng.integers(-5, 6, (4000, 48)).astype(np.float32)
    c = rng.integers(-5, 6, (50, 48)).astype(np.float32)
    c[20] = c[2]
    exact = ((x[:, None, :].astype(np.int64) - c[None, :, :].astype(np.int64)) ** 2).sum(-1)
    order = np.argsort(exact, axis=1, kind="stable")
    assert np.array_equal(qz.closest_centers(x, c, 4), order[:, :4])


def test_closest_centers_errors():
    x = np.zeros((10, 8), np.float32)
    with pytest.raises(dk.ANNError):
        qz.closest_centers(x, np.zeros((4, 8), np.float32), 5)     # k > K
    with pytest.raises(dk.ANNError):
        qz.closest_centers(x, np.zeros((40, 8), np.float32), 17)   # k above 16
