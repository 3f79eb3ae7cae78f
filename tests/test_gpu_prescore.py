"""Candidate pre-scoring (dkm_prescore_*): BF16 or FP8 (E4M3) MFMA approximate scores with a
threshold filter, exact FP32 rescore of the best C, top k.

* Every returned distance equals the oracle's f32 distance (the reference arithmetic)
  bit for bit, and results are ordered by (distance, row).
* On small-integer data the bf16 and fp8 scores are exact (E4M3 holds the integers up to
  16, and the power-of-two operand scales are exact), so the best C by approximate score
  are the best C by exact distance and the output must equal the oracle's exhaustive top
  k — including when the threshold filter (and its retries) does the selection, and for
  integers times a power of two far outside the E4M3 range (the operand scaling).
* With C >= n every row is rescored: the output is the exhaustive FP32 top k on any data.
* On real-valued clustered data recall@k against FP64 brute force stays >= 0.99."""
import numpy as np
import pytest

from diskann_amd.errors import ANNError
from diskann_amd.prescore import BF16, FP8, PreScorer
from oracle import pyoracle as po
from test_gpu_search import log

pytestmark = pytest.mark.gpu


def exact_topk(metric, x, q, k):
    d = np.array([po.distance(metric, q, x[i]) for i in range(x.shape[0])], np.float32)
    order = np.lexsort((np.arange(x.shape[0]), d))[:k]
    return order.astype(np.uint32), d[order]


def check_rows(metric, x, qs, ids, ds):
    for i in range(qs.shape[0]):
        for j in range(ids.shape[1]):
            want = np.float32(po.distance(metric, qs[i], x[ids[i, j]]))
            assert ds[i, j].tobytes() == want.tobytes(), (i, j)
        key = list(zip(ds[i].tolist(), ids[i].tolist()))
        assert key == sorted(key)


PRECISIONS = [pytest.param(BF16, id="bf16"), pytest.param(FP8, id="fp8")]


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("metric,dim", [(po.INNER_PRODUCT, 96), (po.L2, 37), (po.L2, 200),
                                        (po.L2, 320), (po.INNER_PRODUCT, 500),
                                        (po.INNER_PRODUCT, 768)])
def test_integer_data_equals_exhaustive(metric, dim, precision):
    # padded dims 128 / 256 / 384 / 512 / 768: every operand width of both precisions (for
    # FP8 each width is its own schedule of threshold work between the MFMAs)
    log("prescore integer metric=%d dim=%d precision=%d" % (metric, dim, precision))
    rng = np.random.default_rng(dim + metric)
    n, nq, k, C = 6000, 12, 10, 300
    x = rng.integers(-3, 4, (n, dim)).astype(np.float32)
    qs = rng.integers(-3, 4, (nq, dim)).astype(np.float32)
    ps = PreScorer(x, metric, precision)
    ids, ds, ms = ps.search(qs, k, C)
    check_rows(metric, x, qs, ids, ds)
    for i in range(nq):
        wi, wd = exact_topk(metric, x, qs[i], k)
        assert np.array_equal(ids[i], wi), i
        assert ds[i].tobytes() == wd.tobytes()
    assert len(ms) == 4 and all(m >= 0 for m in ms)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE_NORMALIZED])
def test_all_rows_rescored_equals_exhaustive(metric, precision):
    rng = np.random.default_rng(5)
    n, dim, nq, k = 900, 128, 8, 20
    x = rng.standard_normal((n, dim)).astype(np.float32)
    qs = rng.standard_normal((nq, dim)).astype(np.float32)
    if metric == po.COSINE_NORMALIZED:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    ids, ds, _ = PreScorer(x, metric, precision).search(qs, k, candidates=n)
    for i in range(nq):
        wi, wd = exact_topk(metric, x, qs[i], k)
        assert np.array_equal(ids[i], wi) and ds[i].tobytes() == wd.tobytes(), i


@pytest.mark.parametrize("precision", PRECISIONS)
def test_recall_on_clustered_unit_vectors(precision):
    rng = np.random.default_rng(11)
    n, dim, nq, k, C = 20000, 768, 40, 10, 400
    cen = rng.standard_normal((64, dim)).astype(np.float32)
    x = cen[rng.integers(0, 64, n)] + 0.5 * rng.standard_normal((n, dim)).astype(np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    qs = cen[rng.integers(0, 64, nq)] + 0.5 * rng.standard_normal((nq, dim)).astype(np.float32)
    qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    ids, ds, _ = PreScorer(x, po.INNER_PRODUCT, precision).search(qs, k, C)
    check_rows(po.INNER_PRODUCT, x, qs, ids, ds)
    ip = qs.astype(np.float64) @ x.astype(np.float64).T
    gt = np.argsort(-ip, axis=1, kind="stable")[:, :k]
    hits = sum(len(set(ids[i].tolist()) & set(gt[i].tolist())) for i in range(nq))
    assert hits / (nq * k) >= 0.99


@pytest.mark.parametrize("precision", PRECISIONS)
def test_small_base_between_c_and_one_and_a_half_c(precision):
    # C < n <= 1.5 C: every row passes the filter (with one common key), so the candidates
    # must be every row, not the first C by id; the output is the exhaustive FP32 top k
    log("prescore small base precision=%d" % precision)
    rng = np.random.default_rng(11)
    n, dim, nq, k, C = 1000, 64, 8, 10, 800
    x = rng.standard_normal((n, dim)).astype(np.float32)
    qs = rng.standard_normal((nq, dim)).astype(np.float32)
    ids, ds, _ = PreScorer(x, po.L2, precision).search(qs, k, C)
    check_rows(po.L2, x, qs, ids, ds)
    for i in range(nq):
        wi, wd = exact_topk(po.L2, x, qs[i], k)
        assert np.array_equal(ids[i], wi) and ds[i].tobytes() == wd.tobytes(), i


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("k", [1, 1024, 1100])
def test_large_k_equals_exhaustive(k, precision):
    # the final top k is sorted in LDS up to k = 1024 and by a segmented sort beyond
    log("prescore k=%d precision=%d" % (k, precision))
    rng = np.random.default_rng(k)
    n, dim, nq, C = 6000, 32, 3, 1500
    x = rng.integers(-3, 4, (n, dim)).astype(np.float32)
    qs = rng.integers(-3, 4, (nq, dim)).astype(np.float32)
    ids, ds, _ = PreScorer(x, po.L2, precision).search(qs, k, C)
    check_rows(po.L2, x, qs, ids, ds)
    for i in range(nq):
        wi, wd = exact_topk(po.L2, x, qs[i], k)
        assert np.array_equal(ids[i], wi) and ds[i].tobytes() == wd.tobytes(), i


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("C", [8, 40])
def test_small_candidate_counts(C, precision):
    # few candidates against many rows: the sample threshold is taken at a floor rank and
    # the lists sized for it
    rng = np.random.default_rng(C)
    n, dim, nq, k = 40000, 64, 16, 5
    x = rng.integers(-3, 4, (n, dim)).astype(np.float32)
    qs = rng.integers(-3, 4, (nq, dim)).astype(np.float32)
    ids, ds, _ = PreScorer(x, po.INNER_PRODUCT, precision).search(qs, k, C)
    for i in range(nq):
        wi, wd = exact_topk(po.INNER_PRODUCT, x, qs[i], k)
        assert np.array_equal(ids[i], wi) and ds[i].tobytes() == wd.tobytes(), i


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("dim", [64, 768])
def test_many_queries_span_batches(dim, precision):
    # more queries than one filter launch holds (1024): the batches must not interfere; at
    # 768 dims a full launch has the largest LDS footprint (column tiles, thresholds, pool)
    log("prescore batches dim=%d precision=%d" % (dim, precision))
    rng = np.random.default_rng(3)
    n, nq, k, C = 3000, 1500, 5, 100
    x = rng.integers(-2, 3, (n, dim)).astype(np.float32)
    qs = rng.integers(-2, 3, (nq, dim)).astype(np.float32)
    ids, ds, _ = PreScorer(x, po.L2, precision).search(qs, k, C)
    for i in (0, 1023, 1024, 1499):
        wi, wd = exact_topk(po.L2, x, qs[i], k)
        assert np.array_equal(ids[i], wi) and ds[i].tobytes() == wd.tobytes(), i


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT])
def test_fp8_operand_scaling(metric):
    # integers times powers of two far outside E4M3's range (rows up to 3 * 2^12, queries
    # down to 2^-9 for inner product): the per-matrix power-of-two scales bring them into
    # range exactly, and the f32 distances of such values are exact as well
    rng = np.random.default_rng(17 + metric)
    n, dim, nq, k, C = 5000, 128, 10, 10, 200
    x = (rng.integers(-3, 4, (n, dim)) * 4096.0).astype(np.float32)
    qscale = 4096.0 if metric == po.L2 else 1.0 / 512.0
    qs = (rng.integers(-3, 4, (nq, dim)) * qscale).astype(np.float32)
    ids, ds, _ = PreScorer(x, metric, FP8).search(qs, k, C)
    for i in range(nq):
        wi, wd = exact_topk(metric, x, qs[i], k)
        assert np.array_equal(ids[i], wi) and ds[i].tobytes() == wd.tobytes(), i


def test_errors():
    x = np.zeros((100, 16), np.float32)
    with pytest.raises(ANNError):
        PreScorer(x, po.COSINE)
    with pytest.raises(ANNError):
        PreScorer(np.zeros((10, 800), np.float32), po.L2)
    with pytest.raises(ANNError):
        PreScorer(np.zeros((10, 800), np.float32), po.L2, FP8)
    with pytest.raises(ANNError):
        PreScorer(x, po.L2, 7)
    ps = PreScorer(x, po.L2)
    with pytest.raises(ANNError):
        ps.search(np.zeros((1, 16), np.float32), 0)
    with pytest.raises(ANNError):
        ps.search(np.zeros((1, 16), np.float32), 101)
