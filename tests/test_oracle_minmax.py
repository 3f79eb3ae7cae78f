"""8-bit MinMax quantization on the oracle (diskann-quantization/src/minmax).

Pins (no golden vectors exist for MinMax in the reference):
* vectors on the code grid: integer values 0..255 give codes equal to the values and
  exact metadata (b = 0, a = 1, n = sum, |r|^2 = sum of squares);
* the reference's reconstruction test (minmax/quantizer.rs test_quantizer_encoding_random):
  every in-range value reconstructs within half a step, also with a widened grid;
* distances against the reconstructed vectors a * X' + b in f64 (data x data and full-
  precision query x data, all four metrics): a dropped term fails;
* the transform is the spherical quantizer's (pinned there): compressing through a
  transform equals compressing the transformed vector with the null transform."""
import numpy as np
import pytest

from oracle import pyoracle as po


def null(dim):
    return po.Spherical(po.L2, np.zeros(dim, np.float32), 1.0, 1.0, po.SPH_NULL)


def recon(code_row):
    b, n, a, ns = np.frombuffer(code_row[4:20].tobytes(), np.float32)
    return code_row[20:].astype(np.float64) * np.float64(a) + np.float64(b)


def test_grid_values_exact():
    x = np.arange(256, dtype=np.float32)[np.random.default_rng(0).permutation(256)]
    q = po.MinMax(null(256))
    c, loss = q.compress(x)
    assert np.array_equal(c[0, 20:], x.astype(np.uint8))
    dim = np.frombuffer(c[0, :4].tobytes(), np.uint32)[0]
    b, n, a, ns = np.frombuffer(c[0, 4:20].tobytes(), np.float32)
    assert dim == 256 and b == 0.0 and a == 1.0
    assert n == np.float32(x.sum()) and ns == np.float32((x.astype(np.float64) ** 2).sum())
    assert loss[0] == 0.0


@pytest.mark.parametrize("grid", [1.0, 1.25])
@pytest.mark.parametrize("dim", [7, 64, 300])
def test_reconstruction_within_half_step(grid, dim):
    rng = np.random.default_rng(dim)
    x = rng.uniform(-1, 1, (20, dim)).astype(np.float32)
    q = po.MinMax(null(dim), grid)
    c, loss = q.compress(x)
    for r in range(20):
        a = np.frombuffer(c[r, 12:16].tobytes(), np.float32)[0]
        err = np.abs(recon(c[r]) - x[r])
        assert err.max() <= a / 2 + 1e-6
        assert abs(loss[r] - (err ** 2).sum()) <= 1e-4 * max(1.0, loss[r])


@pytest.mark.parametrize("metric", [po.L2, po.INNER_PRODUCT, po.COSINE, po.COSINE_NORMALIZED])
def test_distances_match_reconstruction(metric):
    rng = np.random.default_rng(metric)
    dim = 96
    x = rng.standard_normal((30, dim)).astype(np.float32)
    q = po.MinMax(null(dim))
    c, _ = q.compress(x)
    for i, j in [(0, 1), (3, 17), (29, 5), (8, 8)]:
        rx, ry = recon(c[i]), recon(c[j])
        ip = rx @ ry
        exp = {po.L2: ((rx - ry) ** 2).sum(), po.INNER_PRODUCT: -ip,
               po.COSINE: 1 - ip / (np.linalg.norm(rx) * np.linalg.norm(ry)),
               po.COSINE_NORMALIZED: 1 - ip}[metric]
        got = q.dist_data(metric, c[i], c[j])
        assert abs(got - exp) <= 1e-4 * (1 + abs(rx @ rx) + abs(ry @ ry)), (i, j, got, exp)
        v, m = q.compress_full(x[i])
        ipf = x[i].astype(np.float64) @ ry
        expf = {po.L2: ((x[i] - ry) ** 2).sum(), po.INNER_PRODUCT: -ipf,
                po.COSINE: 1 - ipf / (np.linalg.norm(x[i]) * np.linalg.norm(ry)),
                po.COSINE_NORMALIZED: 1 - ipf}[metric]
        gotf = q.dist_full(metric, v, m, c[j])
        assert abs(gotf - expf) <= 1e-4 * (1 + abs(ry @ ry) + float(x[i] @ x[i])), (i, j)


def test_transform_composes():
    rng = np.random.default_rng(4)
    dim = 100
    t = po.Spherical(po.L2, np.zeros(dim, np.float32), 1.0, 1.0, po.SPH_DOUBLE_HADAMARD,
                     rng.integers(0, 2, dim).astype(bool), rng.integers(0, 2, dim).astype(bool))
    x = rng.standard_normal((5, dim)).astype(np.float32)
    a, _ = po.MinMax(t).compress(x)
    b, _ = po.MinMax(null(dim)).compress(np.stack([t.transform(r) for r in x]))
    assert np.array_equal(a, b)


def test_nan_is_an_error():
    x = np.ones(16, np.float32)
    x[3] = np.nan
    with pytest.raises(ValueError):
        po.MinMax(null(16)).compress(x)
