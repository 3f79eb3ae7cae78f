"""Oracle pins for the direct PQ distances (fixed_chunk_pq_table.rs:32-56, :290-365;
metric dispatch dynamic.rs:111-135).

* test_quant_quant_distances KAT (dynamic.rs:411-474): seeded integer pivot tables
  (test_utils.rs seed_pivot_table / generate_expected_vector); code-vs-code distances
  equal the full-precision distance of the reconstructed vectors exactly for L2, inner
  product and cosine, and within the KAT's tolerance (relative 4e-6, absolute f32 epsilon)
  for cosine-normalized on normalized vectors.
* The full-precision side on integer queries: equal to the reconstructed distance exactly.
* Float queries N(0, 10) (test_utils.rs TestDistribution): within the reference tests'
  tolerances (dynamic.rs:244-378).
* One chunk spanning all dimensions reduces to the full-precision distance function
  (whose 32-partial order is pinned separately), bit for bit.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

METRICS = [po.L2, po.INNER_PRODUCT, po.COSINE, po.COSINE_NORMALIZED]


def seed_pivot_table(dim, pq_chunks, num_pivots, start=0.0):
    off = po.chunk_offsets(dim, pq_chunks)
    piv = np.zeros((num_pivots, dim), np.float32)
    for i in range(num_pivots):
        for j in range(pq_chunks):
            piv[i, off[j]:off[j + 1]] = start + (i + j)
    return piv, off


def expected_vector(code, off, start=0.0):
    v = []
    for i, c in enumerate(code):
        v += [start + (i + int(c))] * int(off[i + 1] - off[i])
    return np.array(v, np.float32)


@pytest.mark.parametrize("dim,pq_chunks", [(20, 7), (200, 7), (20, 20)])
def test_quant_quant_kat(dim, pq_chunks):
    piv, off = seed_pivot_table(dim, pq_chunks, 20)
    rng = np.random.default_rng(0xd49347d0)
    for _ in range(100):
        c0 = rng.integers(0, 20, pq_chunks).astype(np.uint8)
        c1 = rng.integers(0, 20, pq_chunks).astype(np.uint8)
        v0, v1 = expected_vector(c0, off), expected_vector(c1, off)
        for m in (po.L2, po.INNER_PRODUCT, po.COSINE):
            got = po.pq_qq(m, piv, off, c0, c1)
            assert np.float32(got) == np.float32(po.distance(m, v0, v1)), (m, c0, c1)
        sim = po.pq_qq(po.COSINE, piv, off, c0, c1)
        assert 0.0 <= sim <= 2.0
        # normalize as the KAT does: f32 norm (sum of squares in the f32x8 order), f32 divide
        z = np.zeros(dim, np.float32)
        n0 = v0 / np.sqrt(np.float32(po.distance(po.L2, v0, z)), dtype=np.float32)
        n1 = v1 / np.sqrt(np.float32(po.distance(po.L2, v1, z)), dtype=np.float32)
        # CosineNormalized::evaluate = 1 - (inner product in the f32x8 order)
        # (implementations.rs:368-400, IP schema)
        want = float(np.float32(1.0) - np.float32(-po.distance(po.INNER_PRODUCT, n0, n1)))
        got = po.pq_qq(po.COSINE_NORMALIZED, piv, off, c0, c1)
        # approx::assert_relative_eq: absolute epsilon f32::EPSILON or max_relative 4e-6
        assert abs(got - want) <= max(4e-6 * max(abs(want), abs(got)), 1.1920929e-07)


@pytest.mark.parametrize("dim,pq_chunks,num_pivots", [(50, 8, 10), (51, 19, 200), (128, 5, 256),
                                                      (12, 2, 16), (117, 15, 77)])
def test_full_precision_side_integer_query(dim, pq_chunks, num_pivots):
    piv, off = seed_pivot_table(dim, pq_chunks, num_pivots)
    rng = np.random.default_rng(dim * 1000 + pq_chunks)
    for _ in range(40):
        code = rng.integers(0, num_pivots, pq_chunks).astype(np.uint8)
        q = rng.integers(-20, 21, dim).astype(np.float32)
        v = expected_vector(code, off)
        for m in (po.L2, po.INNER_PRODUCT, po.COSINE):
            got = po.pq_direct(m, piv, off, q, code)
            assert np.float32(got) == np.float32(po.distance(m, q, v)), (m, code)


def test_full_precision_side_float_query_tolerances():
    rng = np.random.default_rng(0x83aa68de)
    for dim in (50, 51):
        for pq_chunks in (8, 19, 50):
            piv, off = seed_pivot_table(dim, pq_chunks, 200)
            for _ in range(30):
                code = rng.integers(0, 200, pq_chunks).astype(np.uint8)
                q = (10.0 * rng.standard_normal(dim)).astype(np.float32)
                v = expected_vector(code, off)
                want = po.distance(po.L2, q, v)
                got = po.pq_direct(po.L2, piv, off, q, code)
                assert abs(got - want) <= 6.3e-7 * abs(want)
    for dim in (12, 15, 128):
        for pq_chunks in (2, 5, 15):
            if pq_chunks > dim:
                continue
            piv, off = seed_pivot_table(dim, pq_chunks, 58)
            for _ in range(30):
                code = rng.integers(0, 58, pq_chunks).astype(np.uint8)
                q = (10.0 * rng.standard_normal(dim)).astype(np.float32)
                v = expected_vector(code, off)
                want = po.distance(po.INNER_PRODUCT, q, v)
                got = po.pq_direct(po.INNER_PRODUCT, piv, off, q, code)
                assert abs(got - want) <= max(6.0e-4 * abs(want), 5.0e-3)
                want = po.distance(po.COSINE, q, v)
                got = po.pq_direct(po.COSINE, piv, off, q, code)
                assert abs(got - want) <= 2.0e-7 * max(abs(want), 1.0)


@pytest.mark.parametrize("dim", [7, 8, 33, 100, 128, 768])
def test_single_chunk_equals_full_distance(dim):
    rng = np.random.default_rng(dim)
    piv = rng.standard_normal((4, dim)).astype(np.float32)
    off = np.array([0, dim], np.int64)
    q = rng.standard_normal(dim).astype(np.float32)
    for m in METRICS:
        # the PQ table computes CosineNormalized with its cosine functions, while the f32
        # CosineNormalized functor is 1 - inner product
        fm = po.COSINE if m == po.COSINE_NORMALIZED else m
        for c in range(4):
            code = np.array([c], np.uint8)
            assert np.float32(po.pq_direct(m, piv, off, q, code)) == \
                np.float32(po.distance(fm, q, piv[c]))
            assert np.float32(po.pq_qq(m, piv, off, np.array([1], np.uint8), code)) == \
                np.float32(po.distance(fm, piv[1], piv[c]))


def test_fp64_bound_random_tables():
    rng = np.random.default_rng(7)
    for dim, M in [(128, 32), (96, 48), (768, 12), (50, 19)]:
        piv = rng.standard_normal((256, dim)).astype(np.float32)
        off = po.chunk_offsets(dim, M)
        for _ in range(20):
            code = rng.integers(0, 256, M).astype(np.uint8)
            q = rng.standard_normal(dim).astype(np.float32)
            v = np.concatenate([piv[code[j], off[j]:off[j + 1]] for j in range(M)])
            d64 = float(((q.astype(np.float64) - v) ** 2).sum())
            assert abs(po.pq_direct(po.L2, piv, off, q, code) - d64) <= 1e-5 * d64
            ip64 = float((q.astype(np.float64) * v).sum())
            assert abs(po.pq_direct(po.INNER_PRODUCT, piv, off, q, code) + ip64) <= 1e-4 * (
                1 + abs(ip64))
