"""GPU medoid and exact FP64 brute force against the oracle."""
import numpy as np
import pytest

import diskann_amd as dk
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,dim", [(1, 8), (777, 33), (5000, 128), (3000, 96)])
def test_medoid_matches_oracle(n, dim):
    rng = np.random.default_rng(n + dim)
    x = (rng.standard_normal((n, dim)) * 3 + 1).astype(np.float32)
    ix = dk.DiskANNIndex(dk.Metric.L2, dim, n, 8, 10)
    ix.set_vectors(x)
    row, vec = ix.medoid()
    orow, _ = po.medoid(x)
    assert row == orow
    assert np.array_equal(vec, x[row])


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE])
def test_bruteforce_f64_matches_oracle(metric):
    rng = np.random.default_rng(metric)
    n, dim, k = 4000, 40, 17
    x = rng.standard_normal((n, dim)).astype(np.float32)
    x[10] = x[11]  # an exact tie: smaller id first
    q = rng.standard_normal((23, dim)).astype(np.float32)
    q[0] = x[10]
    ix = dk.DiskANNIndex(metric, dim, n, 8, 10)
    ix.set_vectors(x)
    gi, gd = ix.bruteforce(q, k)
    oi, od = po.bruteforce_f64(metric, x, q, k)
    assert np.array_equal(gd, od)
    assert np.array_equal(gi, oi)
