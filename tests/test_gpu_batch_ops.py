"""Parity of the HIP batch operations with the oracle (batch/mod.rs:62-336).

Bar: bit-identical (each element summed from +0 in the reference's order,
product rounded before the add), f64 and f32, including ragged shapes that do
not fill the kernels' tiles, k = 0, and a batch large enough for many
workgroups in every grid dimension.
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from scirs_amd import linalg
    return linalg


def test_reference_kats(L):
    a = np.array([1.0, 2, 3, 4, 5, 6, 7, 8]).reshape(2, 2, 2)
    assert L.batch_matmul(a, np.array([[1.0, 2], [3, 4]])).ravel().tolist() == [7, 10, 15, 22, 23, 34, 31, 46]
    assert L.batch_matvec(a, np.array([1.0, 2])).ravel().tolist() == [5, 11, 17, 23]
    assert L.batch_add(a, np.array([10.0, 20]), 1).ravel().tolist() == [11, 22, 13, 24, 15, 26, 17, 28]
    assert L.batch_add(a, np.array([10.0, 20]), 0).ravel().tolist() == [11, 12, 23, 24, 15, 16, 27, 28]
    assert L.batch_sum(a).ravel().tolist() == [6, 8, 10, 12]


SHAPES = [(1, 1, 1, 1), (2, 3, 5, 7), (4, 65, 17, 63), (7, 100, 33, 129), (3, 64, 64, 64), (50, 130, 200, 70),
          (1, 1000, 1500, 300)]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("shape", SHAPES)
def test_matmul_matvec_bit_identical(L, dt, shape):
    nb, m, k, n = shape
    print("batch", dt.__name__, shape, flush=True)
    rng = np.random.default_rng(nb * 1000 + m + k + n)
    a = rng.standard_normal((nb, m, k)).astype(dt)
    b = rng.standard_normal((k, n)).astype(dt)
    x = rng.standard_normal(k).astype(dt)
    assert np.array_equal(L.batch_matmul(a, b), oracle.batch_matmul(a, b))
    assert np.array_equal(L.batch_matvec(a, x), oracle.batch_matvec(a, x))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_add_sum_bit_identical(L, dt):
    rng = np.random.default_rng(8)
    for nb, m, n in [(1, 1, 1), (3, 5, 7), (9, 130, 257), (40, 300, 301)]:
        a = rng.standard_normal((nb, m, n)).astype(dt)
        vr = rng.standard_normal(m).astype(dt)
        vc = rng.standard_normal(n).astype(dt)
        assert np.array_equal(L.batch_add(a, vr, 0), oracle.batch_add(a, vr, 0))
        assert np.array_equal(L.batch_add(a, vc, 1), oracle.batch_add(a, vc, 1))
        assert np.array_equal(L.batch_sum(a), oracle.batch_sum(a))


def test_signed_zero_and_empty_inner_dimension(L):
    # all products -0.0: the reference's sum from +0 gives +0 (no padding term may change it)
    a = np.full((2, 3, 5), -0.0)
    b = np.ones((5, 4))
    c = L.batch_matmul(a, b)
    assert np.array_equal(np.signbit(c), np.signbit(oracle.batch_matmul(a, b)))
    assert np.array_equal(L.batch_matmul(np.ones((2, 3, 0)), np.ones((0, 4))), np.zeros((2, 3, 4)))
    assert L.batch_matmul(np.ones((0, 3, 2)), np.ones((2, 4))).shape == (0, 3, 4)


def test_invalid_axis_is_rejected(L):
    with pytest.raises(L.GpuSolveError):
        L.batch_add(np.ones((2, 2, 2)), np.ones(2), 2)
