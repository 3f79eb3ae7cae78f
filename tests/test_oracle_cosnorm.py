"""Pin the oracle's f32 CosineNormalized distance: the CosineNormalized functor runs the
inner-product loop (architecture_hook!(CosineNormalized, simd::IP),
diskann-vector/src/distance/implementations.rs:400) and applies cosine_transformation,
1 - x (implementations.rs:330-332, 368-380); the f32 DistanceProvider builds that functor
(distance_provider.rs:266-270). It is not the full cosine, which the integer element types
use instead (distance_provider.rs:274-299)."""
import numpy as np

from oracle import pyoracle as po


def _f32(x):
    return np.float32(x)


def test_cosnorm_is_one_minus_inner_product_bitwise():
    rng = np.random.default_rng(11)
    for dim in (1, 7, 8, 31, 96, 100, 128, 768, 1536):
        for _ in range(6):
            q = rng.standard_normal(dim).astype(np.float32)
            x = rng.standard_normal(dim).astype(np.float32)
            ip = _f32(po.distance(po.INNER_PRODUCT, q, x))  # = -(sum q x) in the f32x8 order
            got = _f32(po.distance(po.COSINE_NORMALIZED, q, x))
            assert got == _f32(_f32(1.0) - (-ip))


def test_cosnorm_fp64_bound_and_differs_from_cosine_off_the_sphere():
    rng = np.random.default_rng(12)
    for dim in (16, 96, 768):
        q = (3.0 * rng.standard_normal(dim)).astype(np.float32)
        x = (0.5 * rng.standard_normal(dim)).astype(np.float32)
        d64 = 1.0 - float(np.dot(q.astype(np.float64), x.astype(np.float64)))
        got = float(po.distance(po.COSINE_NORMALIZED, q, x))
        scale = float(np.abs(q.astype(np.float64) * x).sum()) + 1.0
        assert abs(got - d64) <= 1e-6 * scale
        # not the full cosine: the vectors are not unit norm
        cos = float(po.distance(po.COSINE, q, x))
        assert abs(cos - got) > 1e-3


def test_cosnorm_equals_cosine_on_unit_vectors_within_rounding():
    rng = np.random.default_rng(13)
    for dim in (96, 128, 1536):
        q = rng.standard_normal(dim)
        x = rng.standard_normal(dim)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        x = (x / np.linalg.norm(x)).astype(np.float32)
        a = float(po.distance(po.COSINE_NORMALIZED, q, x))
        b = float(po.distance(po.COSINE, q, x))
        assert abs(a - b) <= 1e-5


def test_closed_forms():
    e0 = np.zeros(8, np.float32)
    e0[0] = 1.0
    e1 = np.zeros(8, np.float32)
    e1[1] = 1.0
    assert po.distance(po.COSINE_NORMALIZED, e0, e0) == 0.0
    assert po.distance(po.COSINE_NORMALIZED, e0, e1) == 1.0
    assert po.distance(po.COSINE_NORMALIZED, e0, -e0) == 2.0
    # zero vector: no subnormal guard, the inner product is 0 so the distance is 1
    assert po.distance(po.COSINE_NORMALIZED, e0, np.zeros(8, np.float32)) == 1.0


def test_fp64_bruteforce_cosnorm_is_one_minus_dot():
    rng = np.random.default_rng(14)
    base = rng.standard_normal((200, 24)).astype(np.float32)
    qs = rng.standard_normal((3, 24)).astype(np.float32)
    ids, ds = po.bruteforce_f64(po.COSINE_NORMALIZED, base, qs, 5)
    for i in range(3):
        d = 1.0 - base.astype(np.float64) @ qs[i].astype(np.float64)
        order = np.lexsort((np.arange(200), d))[:5]
        assert list(ids[i]) == list(order)
        assert np.allclose(ds[i], d[order], rtol=0, atol=1e-12)
