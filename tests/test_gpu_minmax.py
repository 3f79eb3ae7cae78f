"""GPU 8-bit MinMax quantization against the oracle: code rows (codes and metadata),
reconstruction loss, data x data and full-precision query x data distances for every
metric and transform - bit for bit."""
import numpy as np
import pytest

import diskann_amd as dk
from diskann_amd import minmax
from oracle import pyoracle as po
from test_gpu_search import log

pytestmark = pytest.mark.gpu

CASES = [("null", 37, None, 1.0), ("null", 128, None, 1.25), ("double_hadamard", 100, None, 1.0),
         ("double_hadamard", 200, 150, 1.0), ("padding_hadamard", 768, None, 1.1),
         ("double_hadamard", 1536, None, 1.0)]
KIND = {"null": po.SPH_NULL, "padding_hadamard": po.SPH_PADDING_HADAMARD,
        "double_hadamard": po.SPH_DOUBLE_HADAMARD}


def pair(kind, dim, sub, grid, seed):
    rng = np.random.default_rng(seed)
    s0 = rng.integers(0, 2, dim).astype(bool)
    padded = 1 << (dim - 1).bit_length()
    s1 = subs = None
    if kind == "double_hadamard":
        s1 = rng.integers(0, 2, dim).astype(bool)
        subs = None if sub is None else np.sort(rng.choice(dim, sub, replace=False))
    t = po.Spherical(po.L2, np.zeros(dim, np.float32), 1.0, 1.0, KIND[kind],
                     s0 if kind != "null" else None, s1,
                     padded if kind == "padding_hadamard" else 0, subs)
    g = minmax.MinMaxQuantizer(dim, grid, kind, s0 if kind != "null" else None, s1,
                               padded if kind == "padding_hadamard" else 0, subs)
    return po.MinMax(t, grid), g


@pytest.mark.parametrize("kind,dim,sub,grid", CASES)
def test_minmax_matches_oracle(kind, dim, sub, grid):
    log("minmax %s dim=%d sub=%s grid=%g" % (kind, dim, sub, grid))
    o, g = pair(kind, dim, sub, grid, dim)
    rng = np.random.default_rng(dim + 1)
    x = (rng.standard_normal((50, dim)) * rng.uniform(0.5, 2, dim)).astype(np.float32)
    codes, loss = g.compress(x, with_loss=True)
    oc, ol = o.compress(x)
    assert np.array_equal(codes, oc)
    assert loss.tobytes() == ol.tobytes()
    ids = rng.integers(0, 50, (8, 11)).astype(np.uint32)
    for metric in (po.L2, po.INNER_PRODUCT, po.COSINE, po.COSINE_NORMALIZED):
        dd = g.distance(metric, codes[:8], codes, ids)
        qd = g.query_distance(metric, x[:8] * np.float32(0.9), codes, ids)
        for i in range(8):
            v, m = o.compress_full(x[i] * np.float32(0.9))
            for j in range(11):
                r = ids[i, j]
                assert dd[i, j] == np.float32(o.dist_data(metric, codes[i], codes[r])), (metric, i, j)
                assert qd[i, j] == np.float32(o.dist_full(metric, v, m, codes[r])), (metric, i, j)


def test_minmax_errors():
    g = minmax.MinMaxQuantizer(16)
    x = np.ones((2, 16), np.float32)
    x[1, 2] = np.nan
    with pytest.raises(dk.ANNError):
        g.compress(x)
    with pytest.raises(dk.ANNError):
        minmax.MinMaxQuantizer(16, grid_scale=0.0)
    c = g.compress(np.ones((2, 16), np.float32))
    with pytest.raises(dk.ANNError):
        g.distance(po.L2, c[:1], c, np.array([[5]], np.uint32))
