"""Parity of the HIP Householder QR and least-squares solve with the oracle.

Bars (floating point; tolerances written in each test):
* Q and R equal to the oracle's (lapack.rs:171-253 restated) within
  64 * rows * 2^-53 * max(1, ||A||_max) entry by entry — same reflector
  sequence and signs, sums in a different order;
* at sizes the oracle would take long on: ||Q^T Q - I||_max and
  ||Q R - A||_max / ||A||_max within 64 * rows * 2^-53, R upper triangular to
  that level;
* lstsq: x within 1e-9 relative of the oracle's (solve.rs:234-290), the
  residual sum of squares within 1e-9 relative; the breakdown row of a
  rank-deficient R equal to the oracle's.
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

U64 = 2.0 ** -53


@pytest.fixture(scope="module")
def L():
    from scirs_amd import linalg
    return linalg


@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (3, 2), (5, 3), (33, 33), (64, 64), (200, 77), (513, 300)])
def test_qr_matches_oracle(L, shape):
    print("qr", shape, flush=True)
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    a = rng.standard_normal(shape)
    q, r = L.qr(a)
    qo, ro = oracle.qr_factor(a)
    tol = 64 * shape[0] * U64 * max(1.0, np.abs(a).max())
    assert np.abs(q - qo).max() <= tol
    assert np.abs(r - ro).max() <= tol


def test_qr_reference_examples(L):
    for a in ([[2.0, 1.0], [4.0, 3.0]], [[1.0, 2.0], [3.0, 4.0]], [[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]]):
        a = np.array(a)
        q, r = L.qr(a)
        qo, ro = oracle.qr_factor(a)
        assert np.allclose(q, qo, atol=1e-14) and np.allclose(r, ro, atol=1e-14)
        assert np.allclose(q.T @ q, np.eye(a.shape[0]), atol=1e-12)
        assert np.allclose(q @ r, a, atol=1e-12)


def test_qr_zero_column_is_skipped(L):
    a = np.array([[0.0, 1.0], [0.0, 2.0], [0.0, 3.0]])
    q, r = L.qr(a)
    qo, ro = oracle.qr_factor(a)
    assert np.all(r[:, 0] == 0.0)
    assert np.allclose(q, qo, atol=1e-15) and np.allclose(r, ro, atol=1e-15)


@pytest.mark.parametrize("shape", [(1500, 1500), (3000, 700)])
def test_qr_invariants_large(L, shape):
    print("qr invariants", shape, flush=True)
    rng = np.random.default_rng(11)
    a = rng.standard_normal(shape)
    q, r = L.qr(a)
    tol = 64 * shape[0] * U64
    assert np.abs(q.T @ q - np.eye(shape[0])).max() <= tol
    assert np.abs(q @ r - a).max() <= tol * np.abs(a).max()
    assert np.abs(np.tril(r, -1)).max() <= tol * np.abs(a).max()


def test_qr_rejects_wide(L):
    with pytest.raises(L.GpuSolveError):
        L.qr(np.ones((2, 3)))


@pytest.mark.parametrize("shape", [(3, 2), (50, 20), (200, 200), (301, 45), (700, 333)])
def test_lstsq_matches_oracle(L, shape):
    print("lstsq", shape, flush=True)
    rng = np.random.default_rng(shape[0] + 7 * shape[1])
    a = rng.standard_normal(shape)
    b = rng.standard_normal(shape[0])
    x, res = L.lstsq(a, b)
    xo, reso = oracle.lstsq(a, b)
    assert np.abs(x - xo).max() <= 1e-9 * max(1.0, np.abs(xo).max())
    assert abs(res - reso) <= 1e-9 * max(1.0, reso)


def test_lstsq_documented_example(L):
    x, res = L.lstsq(np.array([[1.0, 1.0], [1.0, 2.0], [1.0, 3.0]]), np.array([6.0, 9.0, 12.0]))
    assert np.allclose(x, [3.0, 3.0], atol=1e-12) and res < 1e-20


def test_lstsq_rank_deficient_reports_the_row(L):
    a = np.array([[1.0, 0.0, 2.0], [3.0, 0.0, 1.0], [4.0, 0.0, 5.0], [2.0, 0.0, 7.0]])
    with pytest.raises(oracle.OracleError) as eo:
        oracle.lstsq(a, np.ones(4))
    with pytest.raises(L.SingularError) as eg:
        L.lstsq(a, np.ones(4))
    assert eg.value.rc == eo.value.code == 2


def test_lstsq_large_normal_equations(L):
    print("lstsq normal equations 4000 x 1000", flush=True)
    rng = np.random.default_rng(5)
    a = rng.standard_normal((4000, 1000))
    b = rng.standard_normal(4000)
    x, res = L.lstsq(a, b)
    r = b - a @ x
    # optimality: A^T r = 0 to rounding, and the residual it reports
    assert np.abs(a.T @ r).max() <= 1e-9 * np.abs(a).max() * np.abs(b).sum()
    assert abs(res - float(r @ r)) <= 1e-9 * float(r @ r)


def test_blocked_workspace_follows_each_shape(L):
    """The blocked sweep's workspace is kept between calls; a taller matrix
    with fewer columns after a wider one must still get its own layout."""
    rng = np.random.default_rng(21)
    a1 = rng.standard_normal((600, 500))
    q1, r1 = L.qr(a1)
    assert np.abs(q1 @ r1 - a1).max() <= 64 * 600 * U64 * np.abs(a1).max()
    a2 = rng.standard_normal((1300, 90))
    b2 = rng.standard_normal(1300)
    print("lstsq 1300 x 90 after qr 600 x 500", flush=True)
    x, res = L.lstsq(a2, b2)
    xo, reso = oracle.lstsq(a2, b2)
    assert np.abs(x - xo).max() <= 1e-9 * max(1.0, np.abs(xo).max())
    assert abs(res - reso) <= 1e-9 * max(1.0, reso)
