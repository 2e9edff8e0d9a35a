"""Pins the CPU oracle (oracle/ref_linalg.c) before anything is checked against it.

What pins which oracle function:

* lu_factor  — exact known answer of lapack.rs:688-734; pivot sequence equal to
  LAPACK dgetrf's at n = 5..512 (same first-maximum rule) and factors equal to
  rounding; exact rational Gaussian elimination with the same pivot rule on
  small integer matrices; the first-maximum tie rule; the absolute-eps
  singularity rule (lapack.rs:106-120); invariants P*A = L*U, |l_ij| <= 1.
* cholesky   — exact answers of lapack.rs:737-754 and the textbook 3x3 of
  tests/scipy_validation.rs:131-152; LAPACK dpotrf at n = 5..64; exact
  rational Cholesky of perfect-square cases; reads only the lower triangle;
  the a_ii - sum <= 0 rule.
* solve_triangular — the four cases of solve.rs:474-507 and exact rationals.
* solve / solve_multiple — solve.rs:458-472, scipy_validation.rs:178-189,
  residual cases of scipy_compat_*; LAPACK dgetrs at n = 5..512; exact
  rational solutions of integer systems; the n <= 4 inverse branch.

The mutation tests at the end plant plausible mistakes in a pure-Python copy
of the elimination (not in the oracle itself) and show that these checks
reject each of them, i.e. the checks are not a restatement of the formula.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "reference_kats.json")) as f:
    KATS = json.load(f)

EPS64 = np.finfo(np.float64).eps


def _lapack(n):
    return np.load(os.path.join(GOLD, f"lapack_n{n}.npz"))


# ---------------------------------------------------------------- checks ---
def check_lu_kat(lu_factor):
    k = KATS["lu_factor_2x2"]
    lu, piv, _ = lu_factor(np.array(k["a"]))
    assert np.array_equal(lu, np.array(k["lu"]))
    assert list(piv) == k["piv"]


def check_lu_lapack(lu_factor, n):
    d = _lapack(n)
    lu, piv, _ = lu_factor(d["a"])
    assert np.array_equal(piv, d["perm"]), "pivot sequence differs from LAPACK dgetrf"
    if "lu" in d:
        assert np.allclose(lu, d["lu"], rtol=1e-11, atol=1e-11 * n)


def check_lu_invariants(lu_factor, a):
    lu, piv, _ = lu_factor(a)
    n = a.shape[0]
    l = np.tril(lu, -1) + np.eye(n)
    u = np.triu(lu)
    pa = a[piv]
    assert np.all(np.abs(np.tril(lu, -1)) <= 1.0)
    res = np.abs(pa - l @ u).max() / (np.abs(a).max() * n)
    assert res < 10 * EPS64


def exact_lu(a_int):
    """Exact rational GE with partial pivoting, first-max rule."""
    a = [[Fraction(v) for v in row] for row in a_int]
    n = len(a)
    piv = list(range(n))
    for k in range(n):
        p = max(range(k, n), key=lambda i: (abs(a[i][k]), -i))
        a[k], a[p] = a[p], a[k]
        piv[k], piv[p] = piv[p], piv[k]
        for i in range(k + 1, n):
            a[i][k] = a[i][k] / a[k][k]
            for j in range(k + 1, n):
                a[i][j] -= a[i][k] * a[k][j]
    return a, piv


def check_lu_exact(lu_factor, seed):
    rng = np.random.default_rng(seed)
    n = 6
    a = rng.integers(-9, 10, size=(n, n)).astype(np.float64)
    a += np.eye(n) * 0.5  # keeps it non-singular and avoids exact ties
    ex, piv_ex = exact_lu(a.tolist())
    lu, piv, _ = lu_factor(a)
    assert list(piv) == piv_ex
    ex_f = np.array([[float(v) for v in row] for row in ex])
    assert np.allclose(lu, ex_f, rtol=1e-13, atol=1e-13)


def check_lu_tie_rule(lu_factor):
    # column 0 has |1| at rows 0 and 2 and |-1| at row 3: the first maximum (row 0) wins
    a = np.array([[1.0, 2.0, 0.0, 1.0],
                  [0.5, 1.0, 3.0, 0.0],
                  [1.0, 0.0, 1.0, 2.0],
                  [-1.0, 1.0, 1.0, 1.0]])
    _, piv, _ = lu_factor(a)
    assert piv[0] == 0


# ------------------------------------------------------------- the tests ---
def test_lu_kat():
    check_lu_kat(oracle.lu_factor)


@pytest.mark.parametrize("n", [5, 8, 16, 64, 512])
def test_lu_matches_lapack(n):
    check_lu_lapack(oracle.lu_factor, n)


@pytest.mark.parametrize("n", [5, 8, 16, 64])
def test_lu_invariants(n):
    check_lu_invariants(oracle.lu_factor, _lapack(n)["a"])


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_lu_exact_rational(seed):
    check_lu_exact(oracle.lu_factor, seed)


def test_lu_tie_rule():
    check_lu_tie_rule(oracle.lu_factor)


def test_lu_property_set():
    k = KATS["lu_property_set"]
    for m in k["mats"]:
        a = np.array(m)
        p, l, u = oracle.lu(a)
        assert np.allclose(p @ a, l @ u, atol=k["tol"])
        assert np.allclose(np.diag(l), 1.0)
        assert np.all(np.triu(l, 1) == 0) and np.all(np.tril(u, -1) == 0)


def test_lu_singular_absolute_eps():
    # a column whose largest entry is below 2^-52 is "singular" for the reference
    a = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 1.0], [1.0, 2.0, 7.0]])
    with pytest.raises(oracle.OracleError) as e:
        oracle.lu_factor(a)
    assert e.value.code == 2
    # a perfectly well-conditioned matrix scaled below eps is singular too
    with pytest.raises(oracle.OracleError) as e:
        oracle.lu_factor(np.eye(3) * 1e-17)
    assert e.value.code == 1
    assert e.value.cond_estimate == pytest.approx(1e17)
    # just above eps it is not
    oracle.lu_factor(np.eye(3) * 4 * EPS64)


def test_lu_rectangular():
    rng = np.random.default_rng(5)
    for shape in ((7, 4), (4, 7)):
        a = rng.standard_normal(shape)
        p, l, u = oracle.lu(a)
        assert p.shape == (shape[0], shape[0])
        assert l.shape == (shape[0], min(shape))
        assert u.shape == (min(shape), shape[1])
        assert np.allclose(p @ a, l @ u, atol=1e-13)


def test_cholesky_kats():
    for key in ("cholesky_2x2", "cholesky_3x3_textbook"):
        k = KATS[key]
        l = oracle.cholesky(np.array(k["a"]))
        assert np.array_equal(l, np.array(k["l"]))
    k = KATS["cholesky_spd_set"]
    for m in k["mats"]:
        a = np.array(m)
        l = oracle.cholesky(a)
        assert np.allclose(l @ l.T, a, atol=k["tol"])
        assert np.all(np.triu(l, 1) == 0)


@pytest.mark.parametrize("n", [5, 8, 16, 64])
def test_cholesky_matches_lapack(n):
    d = _lapack(n)
    l = oracle.cholesky(d["spd"])
    assert np.allclose(l, d["lchol"], rtol=1e-12, atol=1e-13)
    x = oracle.cholesky_solve(d["spd"], d["b"])
    assert np.allclose(x, d["xs"], rtol=1e-10, atol=1e-12)


def test_cholesky_reads_lower_only():
    a = np.array(KATS["cholesky_3x3_textbook"]["a"])
    junk = a.copy()
    junk[np.triu_indices(3, 1)] = 1e300
    assert np.array_equal(oracle.cholesky(junk), oracle.cholesky(a))


def test_cholesky_exact_rational():
    # A = L L^T with small integer L: the factor is recovered exactly
    rng = np.random.default_rng(7)
    l = np.tril(rng.integers(-4, 5, size=(6, 6))).astype(np.float64)
    np.fill_diagonal(l, rng.integers(1, 5, size=6))
    a = l @ l.T
    assert np.array_equal(oracle.cholesky(a), l)


def test_cholesky_not_pd():
    with pytest.raises(oracle.OracleError) as e:
        oracle.cholesky(np.array([[1.0, 2.0], [2.0, 1.0]]))
    assert e.value.kind == "not_positive_definite" and e.value.code == 2
    with pytest.raises(oracle.OracleError) as e:
        oracle.cholesky(np.array([[0.0, 0.0], [0.0, 1.0]]))
    assert e.value.code == 1


def test_solve_triangular_kats():
    for c in KATS["solve_triangular_cases"]["cases"]:
        x = oracle.solve_triangular(np.array(c["a"]), np.array(c["b"]), c["lower"], c["unit"])
        assert np.allclose(x, c["x"], rtol=1e-15, atol=0)


def test_solve_kats():
    k = KATS["solve_cases"]
    for c in k["cases"]:
        x = oracle.solve(np.array(c["a"]), np.array(c["b"]))
        assert np.allclose(x, c["x"], atol=k["tol"])
    for c in KATS["solve_residual_cases"]["cases"]:
        a, b = np.array(c["a"]), np.array(c["b"])
        x = oracle.solve_multiple(a, b)
        assert np.linalg.norm(a @ x - b) < c["tol"]
    k = KATS["singular_solve"]
    with pytest.raises(oracle.OracleError):
        oracle.solve(np.array(k["a"]), np.array(k["b"]))


@pytest.mark.parametrize("n", [5, 8, 16, 64, 512])
def test_solve_matches_lapack(n):
    d = _lapack(n)
    x = oracle.solve(d["a"], d["b"])
    cond = np.linalg.cond(d["a"], np.inf)
    assert np.abs(x - d["x"]).max() / np.abs(d["x"]).max() < 10 * n * EPS64 * cond
    xm = oracle.solve_multiple(d["a"], d["bm"])
    assert np.abs(xm - d["xm"]).max() / np.abs(d["xm"]).max() < 10 * n * EPS64 * cond


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_solve_exact_rational(n):
    rng = np.random.default_rng(100 + n)
    a = rng.integers(-5, 6, size=(n, n)).astype(np.float64) + 7 * np.eye(n)
    xt = rng.integers(-3, 4, size=n).astype(np.float64)
    b = a @ xt  # exact in binary64 (small integers)
    x = oracle.solve(a, b)
    assert np.allclose(x, xt, rtol=0, atol=1e-13)


def test_solve_f32():
    d = _lapack(64)
    x = oracle.solve(d["a"].astype(np.float32), d["b"].astype(np.float32))
    assert x.dtype == np.float32
    assert np.abs(x - d["x"]).max() / np.abs(d["x"]).max() < 1e-3


# --------------------------------------------- planted-mistake mutations ---
def _py_lu_factor(a, mutation=None):
    """Pure-Python elimination used only to plant mistakes in."""
    lu = np.array(a, dtype=np.float64).copy()
    n, m = lu.shape
    piv = np.arange(n)
    for k in range(min(n, m)):
        col = np.abs(lu[k:, k])
        if mutation == "last_max_tie":
            p = k + len(col) - 1 - int(np.argmax(col[::-1]))
        else:
            p = k + int(np.argmax(col))
        if p != k:
            lu[[k, p]] = lu[[p, k]]
            piv[[k, p]] = piv[[p, k]]
        for i in range(k + 1, n):
            lu[i, k] = lu[i, k] / lu[k, k]
            for j in range(k + 1, m):
                if mutation == "sign":
                    lu[i, j] = lu[i, j] + lu[i, k] * lu[k, j]
                elif mutation == "index":
                    lu[i, j] = lu[i, j] - lu[i, k] * lu[k, min(j + 1, m - 1)]
                elif mutation == "transpose":
                    lu[i, j] = lu[i, j] - lu[i, k] * lu[j, k]
                else:
                    lu[i, j] = lu[i, j] - lu[i, k] * lu[k, j]
    return lu, piv, 1.0


@pytest.mark.parametrize("mutation", ["last_max_tie", "sign", "index", "transpose"])
def test_checks_reject_planted_mistakes(mutation):
    def mutant(a):
        return _py_lu_factor(a, mutation)

    checks = [
        check_lu_kat,
        lambda f: check_lu_lapack(f, 16),
        lambda f: check_lu_invariants(f, _lapack(16)["a"]),
        lambda f: check_lu_exact(f, 0),
        check_lu_tie_rule,
    ]
    caught = 0
    for chk in checks:
        try:
            chk(mutant)
        except AssertionError:
            caught += 1
    assert caught >= 1, f"mutation {mutation} passed every check"


def test_unmutated_python_copy_passes():
    f = lambda a: _py_lu_factor(a)  # noqa: E731
    check_lu_kat(f)
    check_lu_lapack(f, 16)
    check_lu_tie_rule(f)


# ------------------------------------------------------------- QR / lstsq ---
# qr_factor is pinned by LAPACK's Householder QR (numpy.linalg.qr, complete
# mode): dgeqrf's reflector sign beta = -sign(alpha)*||x|| (with sign(0) = +)
# is the reference's alpha rule, so Q and R agree entry for entry to rounding,
# signs included (except where LAPACK leaves a column with a zero tail alone
# and the reference reflects it: the last column of a square matrix); a flipped sign, a dropped factor 2 or a transposed Q fails
# it.  The reference's own checks (lapack.rs:757-774, decomposition.rs:794-815,
# tests/scipy_validation.rs:78-99) are the Q^T Q = I / Q R = A / upper-R ones.
# lstsq is pinned by numpy.linalg.lstsq (x and the residual sum of squares)
# and by the documented example of solve.rs:225-233.
@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (3, 2), (5, 3), (40, 25), (64, 64), (130, 7)])
def test_qr_matches_lapack(shape):
    rng = np.random.default_rng(shape[0] * 131 + shape[1])
    a = rng.standard_normal(shape)
    q, r = oracle.qr_factor(a)
    qn, rn = np.linalg.qr(a, mode="complete")
    if shape[0] == shape[1]:
        # the reference also reflects the last 1-element column (v = e1, H = -1,
        # so r_nn changes sign), where LAPACK's dlarfg takes tau = 0
        qn[:, -1] *= -1.0
        rn[-1, :] *= -1.0
    assert q.shape == (shape[0], shape[0]) and r.shape == shape
    scale = max(1.0, np.abs(a).max())
    assert np.abs(q - qn).max() <= 64 * shape[0] * EPS64
    assert np.abs(r - rn).max() <= 64 * shape[0] * EPS64 * scale
    assert np.abs(q @ r - a).max() <= 64 * shape[0] * EPS64 * scale
    assert np.abs(q.T @ q - np.eye(shape[0])).max() <= 64 * shape[0] * EPS64


def test_qr_reference_examples():
    for a in ([[2.0, 1.0], [4.0, 3.0]], [[1.0, 2.0], [3.0, 4.0]], [[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]]):
        a = np.array(a)
        q, r = oracle.qr_factor(a)
        assert np.allclose(q.T @ q, np.eye(a.shape[0]), atol=1e-12)
        assert np.allclose(q @ r, a, atol=1e-12)
        assert np.allclose(np.tril(r, -1), 0.0, atol=1e-12)
    # [[1, 2], [3, 4]]: r_00 = -sqrt(10) (x0 = 1 >= 0 -> alpha = -||x||)
    _, r = oracle.qr_factor(np.array([[1.0, 2.0], [3.0, 4.0]]))
    assert abs(r[0, 0] + np.sqrt(10.0)) < 1e-14
    assert abs(r[1, 1] - 2.0 / np.sqrt(10.0)) < 1e-14


def test_qr_zero_column_is_skipped():
    a = np.array([[0.0, 1.0], [0.0, 2.0], [0.0, 3.0]])
    q, r = oracle.qr_factor(a)
    # column 0 has ||x|| = 0 <= eps: no reflector; column 1 is reflected from row 1
    assert np.all(r[:, 0] == 0.0)
    assert r[0, 1] == 1.0 and q[0, 0] == 1.0
    assert np.allclose(q @ r, a, atol=1e-14)


def test_qr_rejects_wide_and_empty():
    with pytest.raises(oracle.OracleError):
        oracle.qr_factor(np.ones((2, 3)))
    with pytest.raises(oracle.OracleError):
        oracle.qr_factor(np.ones((0, 0)))


@pytest.mark.parametrize("shape", [(3, 2), (50, 20), (200, 200), (301, 45)])
def test_lstsq_matches_lapack(shape):
    rng = np.random.default_rng(shape[0] + 7 * shape[1])
    a = rng.standard_normal(shape)
    b = rng.standard_normal(shape[0])
    x, res = oracle.lstsq(a, b)
    xn, rn, _, _ = np.linalg.lstsq(a, b, rcond=None)
    assert np.abs(x - xn).max() <= 1e-10 * max(1.0, np.abs(xn).max())
    expect = float(rn[0]) if rn.size else float(np.sum((b - a @ xn) ** 2))
    assert abs(res - expect) <= 1e-10 * max(1.0, expect)


def test_lstsq_documented_example():
    x, res = oracle.lstsq(np.array([[1.0, 1.0], [1.0, 2.0], [1.0, 3.0]]), np.array([6.0, 9.0, 12.0]))
    assert np.allclose(x, [3.0, 3.0], atol=1e-12) and res < 1e-24


def test_lstsq_rank_deficient_reports_the_row():
    a = np.array([[1.0, 0.0, 2.0], [3.0, 0.0, 1.0], [4.0, 0.0, 5.0], [2.0, 0.0, 7.0]])
    with pytest.raises(oracle.OracleError) as e:
        oracle.lstsq(a, np.ones(4))
    assert e.value.code == 2   # r_11 = 0; row 3 (bottom) passes, row 2 fails


# ------------------------------------------------------------ batch ops ---
# The batch operations are pinned by the reference's own tests
# (batch/mod.rs:340-495: exact small answers), by exact integer arithmetic
# (products and sums of small integers are exact in f64, so any order gives
# the same bits: a dropped term or a transposed operand changes them), and by
# numpy to rounding on random inputs.
def test_batch_ops_reference_kats():
    a = np.array([1.0, 2, 3, 4, 5, 6, 7, 8]).reshape(2, 2, 2)
    assert oracle.batch_matmul(a, np.array([[1.0, 2], [3, 4]])).ravel().tolist() == [7, 10, 15, 22, 23, 34, 31, 46]
    assert oracle.batch_matvec(a, np.array([1.0, 2])).ravel().tolist() == [5, 11, 17, 23]
    assert oracle.batch_add(a, np.array([10.0, 20]), 1).ravel().tolist() == [11, 22, 13, 24, 15, 26, 17, 28]
    assert oracle.batch_add(a, np.array([10.0, 20]), 0).ravel().tolist() == [11, 12, 23, 24, 15, 16, 27, 28]
    assert oracle.batch_sum(a).ravel().tolist() == [6, 8, 10, 12]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_batch_ops_exact_integers(dt):
    rng = np.random.default_rng(3)
    a = rng.integers(-50, 50, (5, 7, 9)).astype(dt)
    b = rng.integers(-50, 50, (9, 4)).astype(dt)
    x = rng.integers(-50, 50, 9).astype(dt)
    assert np.array_equal(oracle.batch_matmul(a, b), np.einsum("pik,kj->pij", a, b))
    assert np.array_equal(oracle.batch_matvec(a, x), np.einsum("pij,j->pi", a, x))
    assert np.array_equal(oracle.batch_add(a, x, 1), a + x[None, None, :])
    assert np.array_equal(oracle.batch_add(a, x[:7], 0), a + x[:7][None, :, None])
    assert np.array_equal(oracle.batch_sum(a), a.sum(axis=0))


def test_batch_ops_random_to_rounding():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((3, 20, 33))
    b = rng.standard_normal((33, 11))
    x = rng.standard_normal(33)
    assert np.allclose(oracle.batch_matmul(a, b), a @ b, rtol=0, atol=1e-13)
    assert np.allclose(oracle.batch_matvec(a, x), a @ x, rtol=0, atol=1e-13)
    s = np.zeros((20, 33))
    for p in range(3):
        s = s + a[p]
    assert np.array_equal(oracle.batch_sum(a), s)   # batch order, as ndarray's sum_axis


# ---------------------------------------------------------------- gmres ---
# gmres is pinned by the reference's test (solvers/iterative.rs:690-703:
# [[1,2],[3,4]] x = [5,11] -> [1,2], converged), by LAPACK's solution of
# well-conditioned systems (a wrong sign in the Gram-Schmidt or the Givens
# sweep stalls the residual), by the exact-in-n-steps property (an n x n
# system converges within one cycle of n steps), and by the cycle arithmetic
# ((max_iter / restart).max(1) cycles: max_iter = 5 gives 5 iterations, not
# converged).
def test_gmres_reference_example():
    x, it, res, conv = oracle.gmres(np.array([[1.0, 2.0], [3.0, 4.0]]), np.array([5.0, 11.0]))
    assert conv and abs(x[0] - 1.0) < 1e-10 and abs(x[1] - 2.0) < 1e-10 and it == 2


@pytest.mark.parametrize("n", [5, 40, 200])
def test_gmres_matches_lapack(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 3.0 * np.eye(n)
    b = rng.standard_normal(n)
    x, it, res, conv = oracle.gmres(a, b)
    assert conv and res < 1e-10
    assert np.abs(x - np.linalg.solve(a, b)).max() < 1e-9
    assert abs(res - np.linalg.norm(b - a @ x)) < 1e-12


def test_gmres_cycle_arithmetic_and_x0():
    rng = np.random.default_rng(2)
    n = 60
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 1.2 * np.eye(n)
    b = rng.standard_normal(n)
    x, it, res, conv = oracle.gmres(a, b, max_iter=5)
    assert it == 5 and not conv
    xs = np.linalg.solve(a, b)
    x, it, res, conv = oracle.gmres(a, b, x0=xs)
    assert conv and it == 0 and res < 1e-10
    x, it, res, conv = oracle.gmres(a, b, restart=n)
    assert conv and it <= n


# ------------------------------------- right-preconditioned GMRES (pgmres) ---
# The oracle's restatement of gmres (iterative.rs:317-450) with right
# preconditioning, the real counterpart of preconditioners.rs:934-1010 (whose
# body never solves the least-squares problem, so it cannot serve as a check).
def _row_scaled(n, seed):
    """D (I + E) with D log-spaced over 1e-3 .. 1e3: hard for plain GMRES,
    nearly the identity after Jacobi scaling."""
    rng = np.random.default_rng(seed)
    d = np.logspace(-3, 3, n)
    return d[:, None] * (np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)), rng.standard_normal(n)


def test_pgmres_without_preconditioner_is_gmres():
    rng = np.random.default_rng(4)
    n = 80
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 1.5 * np.eye(n)
    b = rng.standard_normal(n)
    x0, it0, r0, c0 = oracle.gmres(a, b, max_iter=300, restart=20)
    x1, it1, r1, c1 = oracle.pgmres(a, b, 0, max_iter=300, restart=20)
    assert (it0, c0) == (it1, c1) and np.array_equal(x0, x1) and r0 == r1


@pytest.mark.parametrize("n", [30, 120])
def test_pgmres_jacobi_matches_lapack_and_beats_plain(n):
    a, b = _row_scaled(n, n)
    x, it, res, conv = oracle.pgmres(a, b, 1, max_iter=2000, tol=1e-10, restart=30)
    xs = np.linalg.solve(a, b)
    assert conv
    # the reported residual is that of A x = b (right preconditioning), not of M^-1 A
    # (to the rounding of a residual: 8 n u (||A|| ||x|| + ||b||))
    rb = 8 * n * 2.0 ** -53 * (np.linalg.norm(a, 2) * np.linalg.norm(x) + np.linalg.norm(b))
    assert abs(res - np.linalg.norm(b - a @ x)) <= rb
    kappa = np.linalg.cond(a)
    assert np.abs(x - xs).max() <= 1e-10 * kappa * max(1.0, np.abs(xs).max())
    # the reference's loop tests convergence only at the end of a cycle, so
    # iteration counts move in whole cycles: short cycles tell them apart
    _, itp, _, cp = oracle.pgmres(a, b, 1, max_iter=5000, tol=1e-10, restart=10)
    _, it0, _, c0 = oracle.gmres(a, b, max_iter=5000, tol=1e-10, restart=10)
    assert cp and (it0 > 2 * itp or not c0), (itp, it0, c0)


def test_pgmres_exact_lu_takes_one_cycle():
    rng = np.random.default_rng(9)
    n = 150
    a = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    lu, piv, _ = oracle.lu_factor(a)
    x, it, res, conv = oracle.pgmres(a, b, 3, lu=lu, perm=piv, max_iter=300, tol=1e-9, restart=30)
    assert conv and it <= 30
    assert np.abs(x - np.linalg.solve(a, b)).max() <= 1e-9 * np.linalg.cond(a)
    # an FP32 LU preconditioner still solves to FP64 accuracy, in far fewer
    # iterations than no preconditioner
    lu32, piv32, _ = oracle.lu_factor(a.astype(np.float32))
    x, it, res, conv = oracle.pgmres(a, b, 3, lu=lu32.astype(np.float64), perm=piv32, max_iter=300, tol=1e-11,
                                     restart=30)
    _, it0, _, c0 = oracle.gmres(a, b, max_iter=300, tol=1e-11, restart=10)
    _, itp, _, cp = oracle.pgmres(a, b, 3, lu=lu32.astype(np.float64), perm=piv32, max_iter=300, tol=1e-11,
                                  restart=10)
    assert conv and cp and (not c0 or itp < it0), (itp, it0, c0)


def _py_pgmres(a, b, minv, restart, max_iter, tol, mutation=None):
    """Pure-Python right-preconditioned GMRES used only to plant mistakes in."""
    n = a.shape[0]
    x = np.zeros(n)
    total = 0
    for _ in range(max(max_iter // restart, 1)):
        r = b - a @ x
        beta = np.linalg.norm(r)
        if beta < tol:
            break
        V = [r / beta]
        H = np.zeros((restart + 1, restart))
        j = 0
        while j < restart and total < max_iter:
            z = V[j] if mutation == "no_prec_in_arnoldi" else minv(V[j])
            w = a @ z
            wo = w.copy()
            for i in range(j + 1):
                H[i, j] = w @ V[i]
                wo -= H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(wo)
            if H[j + 1, j] < np.finfo(float).eps:
                j += 1
                break
            V.append(wo / H[j + 1, j])
            j += 1
            total += 1
        g = np.zeros(j + 1)
        g[0] = beta
        y = np.linalg.lstsq(H[:j + 1, :j], g, rcond=None)[0]
        t = sum(y[i] * V[i] for i in range(j))
        x = x + (t if mutation == "no_prec_in_update" else minv(t))
    return x


@pytest.mark.parametrize("mutation", ["no_prec_in_arnoldi", "no_prec_in_update"])
def test_pgmres_checks_reject_planted_mistakes(mutation):
    """The LAPACK check above catches a preconditioner left out of the Arnoldi
    step or out of the update (the unmutated Python copy passes it)."""
    a, b = _row_scaled(60, 3)
    dinv = 1.0 / np.diag(a)
    xs = np.linalg.solve(a, b)
    ok = _py_pgmres(a, b, lambda v: dinv * v, 30, 600, 1e-10)
    assert np.abs(ok - xs).max() <= 1e-8 * max(1.0, np.abs(xs).max())
    bad = _py_pgmres(a, b, lambda v: dinv * v, 30, 600, 1e-10, mutation)
    assert not np.abs(bad - xs).max() <= 1e-8 * max(1.0, np.abs(xs).max())
