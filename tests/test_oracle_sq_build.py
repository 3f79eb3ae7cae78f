"""Pin the oracle's 8-bit scalar-quantized build (the Quantized strategy over an SQStore:
insert search with the SQ query computer, diskann-providers/src/model/graph/provider/
async_/inmem/scalar.rs:595-672, glue.rs:933-940; prune and back-edge distances between
stored codes, scalar.rs:370-446, 676-720) by what the arithmetic fixes:

* integer vectors in [0, 255] with shift 0 and a scale whose (1/255)^2 scale^2 factor rounds
  to exactly 1.0: every code equals its value, every compensation is 0 and every SQ L2
  distance (query x data and data x data) is the exact integer squared distance, which the
  f32 full-precision distance also is. The SQ build must then equal the full-precision
  build node for node (single inserts as the reference's SQ benchmark runs them, and
  batches), and so must robust_prune_list;
* on general data the SQ and full-precision builds differ (the SQ path is taken)."""
import numpy as np
import pytest

from oracle import pyoracle as po

# (1/255)^2 * scale^2 rounds to exactly 1.0f at this scale; codes of 0..255 stay exact
GRID_SCALE = float(np.float32(254.99998474121094))


def grid_data(n, dim, seed):
    rng = np.random.default_rng(seed)
    cen = rng.integers(20, 236, (8, dim))
    x = cen[rng.integers(0, 8, n)] + rng.integers(-20, 21, (n, dim))
    return np.clip(x, 0, 255).astype(np.float32)


def sq_of(x, scale, shift, mean_norm=1.0):
    codes, comp = po.sq_compress(scale, shift, x)
    return (scale, shift, mean_norm, codes, comp)


def test_grid_scale_makes_sq_l2_exact():
    x = grid_data(40, 24, 3)
    shift = np.zeros(24, np.float32)
    _, _, _, codes, comp = sq_of(x, GRID_SCALE, shift)
    assert np.array_equal(codes.astype(np.float32), x) and not comp.any()
    ssn = po.sq_shift_square_norm(shift)
    for i in range(0, 40, 3):
        for j in range(40):
            assert po.sq_distance(po.L2, GRID_SCALE, ssn, codes[i], comp[i], codes[j],
                                  comp[j]) == po.distance(po.L2, x[i], x[j])


@pytest.mark.parametrize("batch,intra", [(1, -1), (16, -1), (16, 0)])
def test_sq_build_on_the_code_grid_equals_full_precision(batch, intra):
    n, dim, R = 600, 24, 12
    x = grid_data(n, dim, 5)
    med, _ = po.medoid(x)
    vec = np.vstack([x, x[med:med + 1]])
    sq = sq_of(vec, GRID_SCALE, np.zeros(dim, np.float32))
    kw = dict(R=R, R_max=16, L_build=32, alpha=1.2, starts=[n], ids=np.arange(n), batch=batch,
              intra_batch=intra)
    a_fp, d_fp = po.build_graph(po.L2, vec, **kw)
    a_sq, d_sq = po.build_graph(po.L2, vec, sq=sq, **kw)
    assert np.array_equal(d_fp, d_sq)
    for i in range(n + 1):
        assert list(a_fp[i, :d_fp[i]]) == list(a_sq[i, :d_sq[i]]), i


def test_sq_prune_list_on_the_code_grid_equals_full_precision():
    x = grid_data(300, 16, 9)
    sq = sq_of(x, GRID_SCALE, np.zeros(16, np.float32))
    rng = np.random.default_rng(2)
    for loc in range(0, 300, 37):
        lst = rng.choice(300, 120, replace=False).astype(np.uint32)
        want = po.prune_list(po.L2, x, loc, lst, 10, 1.2)
        got = po.prune_list(po.L2, x, loc, lst, 10, 1.2, sq=sq)
        assert list(got) == list(want)


def test_sq_build_on_general_data_differs_from_full_precision():
    rng = np.random.default_rng(7)
    n, dim = 500, 24
    x = rng.standard_normal((n, dim)).astype(np.float32)
    med, _ = po.medoid(x)
    vec = np.vstack([x, x[med:med + 1]])
    scale, shift, mean_norm = po.sq_train(vec, 2.0)
    sq = sq_of(vec, scale, shift, mean_norm)
    kw = dict(R=12, R_max=16, L_build=32, alpha=1.2, starts=[n], ids=np.arange(n), batch=1)
    a_fp, d_fp = po.build_graph(po.L2, vec, **kw)
    a_sq, d_sq = po.build_graph(po.L2, vec, sq=sq, **kw)
    assert any(list(a_fp[i, :d_fp[i]]) != list(a_sq[i, :d_sq[i]]) for i in range(n))
