"""The C++ mirror of the scirs2-linalg API (include/scirs2_linalg.hpp), driven
through tests/cpp/mirror_driver.

CPU tests cover everything that needs no device: the validation order and the
exact error texts (scirs2-linalg/src/validation.rs:22-329, error.rs:73-150,
decomposition.rs:114-127, basic.rs:38-43/126-131), the closed forms of det for
n <= 3 and of inv for n = 2 (basic.rs:46-56, 134-165), and `solve` for n <= 2,
which the reference routes through inv (solve.rs:65-76).  GPU tests compare the
mirror's results with the oracle (FP64: normwise backward error <= 8 n 2^-53,
pivots and breakdowns identical).
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "mirror_driver")
U64 = 2.0 ** -53
MUL = "×"


def _fmt(v):
    v = float(v)
    if math.isnan(v):
        return "nan"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    return repr(v)


def run(cases):
    """cases: (op, dtype, flag1, flag2, A, B) -> list of ('OK', [arrays]) / ('ERR', kind, text)."""
    src = [os.path.join(ROOT, "tests", "cpp", "mirror_driver.cpp"), os.path.join(ROOT, "include", "scirs2_linalg.hpp")]
    if not os.path.exists(DRIVER) or max(map(os.path.getmtime, src)) > os.path.getmtime(DRIVER):
        from tests.cpp import build_driver   # host-only g++ build, seconds
        build_driver.build()
    lines = []
    for op, dt, f1, f2, a, b in cases:
        a = np.asarray(a, dtype=np.float64).reshape(np.shape(a) if np.ndim(a) == 2 else (0, 0))
        toks = [op, dt, str(int(f1)), str(int(f2)), str(a.shape[0]), str(a.shape[1])]
        toks += [_fmt(v) for v in a.ravel()]
        if b is None:
            toks += ["0", "0"]
        else:
            b = np.asarray(b, dtype=np.float64)
            b2 = b.reshape(-1, 1) if b.ndim == 1 else b
            toks += [str(b2.shape[0]), str(b2.shape[1])] + [_fmt(v) for v in b2.ravel()]
        lines.append(" ".join(toks))
    r = subprocess.run([DRIVER], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = []
    for line in r.stdout.splitlines():
        if line.startswith("ERR "):
            _, kind, text = line.split(" ", 2)
            out.append(("ERR", kind, text.replace("<LF>", "\n")))
            continue
        t = line.split()
        assert t[0] == "OK"
        cnt, pos, arrs = int(t[1]), 2, []
        for _ in range(cnt):
            rr, cc = int(t[pos]), int(t[pos + 1])
            pos += 2
            arrs.append(np.array([float(x) for x in t[pos:pos + rr * cc]]).reshape(rr, cc))
            pos += rr * cc
        out.append(("OK", arrs))
    assert len(out) == len(cases), r.stdout
    return out


# Expected texts, written from the reference's format strings.  The
# validation helpers' strings hold a literal backslash-n (validation.rs:75,
# 104, 134); det / inv and the *_with_suggestions builders use real newlines.
def shape_square(op, r, c):   # validation.rs:73-77
    return f"Shape error: {op} failed: Matrix must be square\\nMatrix shape: {r}{MUL}{c}\\nExpected: Square matrix (n{MUL}n)"


def singular(operation, r, c, cond=None):   # error.rs:73-110
    s = ["Consider the following regularization approaches:",
         "1. Ridge regularization: Add λI to the matrix (small positive λ)",
         "2. Pseudo-inverse: Use SVD-based pseudo-inverse for rank-deficient matrices",
         "3. Truncated SVD: Remove small singular values below a threshold"]
    if cond is not None:
        s.append(f"4. Condition number: {cond} (>1e12 indicates ill-conditioning)")
        if float(cond) > 1e12:
            s.append("5. Use extended precision arithmetic for better numerical stability")
    s += ["6. Check input data for linear dependencies or scaling issues",
          "7. Use iterative refinement for improved accuracy"]
    return (f"Singular matrix error: Matrix is singular during {operation} operation\n"
            f"Matrix shape: {r}{MUL}{c}\n" + "\n".join(s))


def not_pd(operation, r, c):   # error.rs:112-150
    s = ["Consider the following regularization approaches:",
         "1. Diagonal regularization: Add λI where λ > |most negative eigenvalue|",
         "2. Modified Cholesky: Use algorithms that ensure positive definiteness",
         "3. Eigenvalue clipping: Replace negative eigenvalues with small positive values",
         "4. Use LDL decomposition instead of Cholesky for indefinite matrices",
         "6. Check if matrix is symmetric (required for Cholesky)",
         "7. Use pivoted Cholesky for rank-deficient positive semidefinite matrices",
         "8. Consider using QR or LU decomposition for non-symmetric matrices"]
    return (f"Non-positive definite matrix error: Matrix is not positive definite during {operation} operation\n"
            f"Matrix shape: {r}{MUL}{c}\n" + "\n".join(s))


NAN = float("nan")
INF = float("inf")
I3 = np.eye(3)


def test_validation_order_and_texts():
    sq23 = np.arange(6.0).reshape(2, 3)
    cases = [
        # solve: empty matrix before empty vector (validation.rs:220-221)
        ("solve", "f64", 0, 0, np.zeros((0, 0)), np.zeros(0)),
        ("solve", "f64", 0, 0, I3, np.zeros(0)),
        # non-square before non-finite (validation.rs:222-224)
        ("solve", "f64", 0, 0, np.full((2, 3), NAN), [1.0, 2.0]),
        ("solve", "f64", 0, 0, I3, [1.0, 2.0]),
        ("solve", "f64", 0, 0, [[1.0, INF, 0], [0, 1, 0], [0, 0, 1]], [1.0, 2.0, NAN]),
        ("solve", "f64", 0, 0, I3, [1.0, -INF, 3.0]),
        ("solve_multiple", "f64", 0, 0, I3, np.ones((2, 2))),
        ("solve_multiple", "f64", 0, 0, I3, np.full((3, 1), NAN)),
        ("solve_multiple", "f64", 0, 0, sq23, np.ones((2, 2))),
        ("lu", "f64", 0, 0, np.zeros((0, 0)), None),
        ("lu", "f64", 0, 0, [[1.0, NAN, 2.0]], None),      # lu accepts non-square; finite check only
        ("cholesky", "f64", 0, 0, sq23, None),
        ("cholesky", "f64", 0, 0, [[1.0, NAN], [NAN, 1.0]], None),
        ("solve_triangular", "f64", 1, 0, sq23, [1.0, 2.0]),
        ("solve_triangular", "f32", 0, 0, I3, [1.0, 2.0]),
        ("det", "f64", 0, 0, sq23, None),
        ("inv", "f64", 0, 0, sq23, None),
    ]
    got = run(cases)
    mv = (f"Shape error: Linear system solve failed: Matrix and vector dimensions must match\\nMatrix shape: 3{MUL}3"
          f"\\nVector shape: 2\\nExpected: Vector length = 3")
    rows = (f"Shape error: Multiple linear systems solve failed: Matrix rows must match\\nFirst matrix shape: 3{MUL}3"
            f"\\nSecond matrix shape: 2{MUL}2\\nExpected: Same number of rows")
    tri_mv = (f"Shape error: Triangular system solve failed: Matrix and vector dimensions must match\\nMatrix shape: "
              f"3{MUL}3\\nVector shape: 2\\nExpected: Vector length = 3")
    expected = [
        ("ShapeError", "Shape error: Linear system solve failed: Input matrix cannot be empty"),
        ("ShapeError", "Shape error: Linear system solve failed: Input vector cannot be empty"),
        ("ShapeError", shape_square("Linear system solve", 2, 3)),
        ("ShapeError", mv),
        ("InvalidInputError", "Invalid input error: Linear system solve failed: Matrix contains non-finite values"),
        ("InvalidInputError", "Invalid input error: Linear system solve failed: Vector contains non-finite values"),
        ("ShapeError", rows),
        ("InvalidInputError",
         "Invalid input error: Multiple linear systems solve failed: Matrix contains non-finite values"),
        ("ShapeError", shape_square("Multiple linear systems solve", 2, 3)),
        ("ShapeError", "Shape error: LU decomposition failed: Input matrix cannot be empty"),
        ("InvalidInputError", "Invalid input error: LU decomposition failed: Matrix contains non-finite values"),
        ("ShapeError", shape_square("Cholesky decomposition", 2, 3)),
        ("InvalidInputError", "Invalid input error: Cholesky decomposition failed: Matrix contains non-finite values"),
        ("ShapeError", shape_square("Triangular system solve", 2, 3)),
        ("ShapeError", tri_mv),
        ("ShapeError", f"Shape error: Determinant computation failed: Matrix must be square\nMatrix shape: 2{MUL}3"
                       f"\nExpected: Square matrix (n{MUL}n)"),
        ("ShapeError", f"Shape error: Matrix inverse computation failed: Matrix must be square\nMatrix shape: 2{MUL}3"
                       f"\nExpected: Square matrix (n{MUL}n)"),
    ]
    for (case, g, e) in zip(cases, got, expected):
        assert g[0] == "ERR", (case[0], g)
        assert (g[1], g[2]) == e, case[0]


def test_triangular_singular_diagonal_texts():
    # solve.rs:166-173 / 186-193: first |a_ii| < F::epsilon() in substitution order, cond 1e16
    lo = np.array([[2.0, 0, 0], [1, 0.0, 0], [1, 1, 0.0]])
    up = np.array([[0.0, 1, 1], [0, 3, 1], [0, 0, 1e-17]])
    f32_small = np.array([[1.0, 0], [0, 1e-8]])   # below f32 eps, not a breakdown in f64 arithmetic
    got = run([("solve_triangular", "f64", 1, 0, lo, [1.0, 1, 1]),
               ("solve_triangular", "f64", 0, 0, up, [1.0, 1, 1]),
               ("solve_triangular", "f32", 1, 0, f32_small, [1.0, 1])])
    assert got[0] == ("ERR", "SingularMatrixError",
                      singular("triangular system solve (forward substitution)", 3, 3, "1.00e16"))
    assert got[1] == ("ERR", "SingularMatrixError",
                      singular("triangular system solve (back substitution)", 3, 3, "1.00e16"))
    assert got[2] == ("ERR", "SingularMatrixError",
                      singular("triangular system solve (forward substitution)", 2, 2, "1.00e16"))


def test_closed_forms_det_inv_and_small_solve():
    rng = np.random.default_rng(3)
    a1 = rng.standard_normal((1, 1))
    a2 = rng.standard_normal((2, 2))
    a3 = rng.standard_normal((3, 3))
    b2 = rng.standard_normal(2)
    got = run([("det", "f64", 0, 0, np.zeros((0, 0)), None),
               ("det", "f64", 0, 0, a1, None),
               ("det", "f64", 0, 0, a2, None),
               ("det", "f64", 0, 0, a3, None),
               ("inv", "f64", 0, 0, a2, None),
               ("solve", "f64", 0, 0, a2, b2),
               ("solve", "f32", 0, 0, a2, b2)])
    assert got[0][1][0][0, 0] == 1.0   # basic.rs:47: det of 0x0 is one
    assert got[1][1][0][0, 0] == a1[0, 0]
    for k, a in ((2, a2), (3, a3)):
        assert abs(got[k][1][0][0, 0] - np.linalg.det(a)) <= 1e-13 * max(1.0, abs(np.linalg.det(a)))
    inv2 = got[4][1][0]
    assert np.allclose(inv2 @ a2, np.eye(2), atol=1e-13)
    # solve for n <= 4 is the product with the inverse (solve.rs:65-76): same as the oracle
    xo = oracle.solve(a2, b2)
    assert np.allclose(got[5][1][0].ravel(), xo, rtol=0, atol=1e-14 * np.abs(xo).max())
    xo32 = oracle.solve(a2.astype(np.float32), b2.astype(np.float32))
    assert np.allclose(got[6][1][0].ravel(), xo32, rtol=0, atol=1e-6 * np.abs(xo32).max())


def test_inverse_2x2_singular_condition_line():
    # basic.rs:136-157: |det| < eps; cond = ||A||_F / |det| only when det != 0
    got = run([("inv", "f64", 0, 0, [[1e-9, 0], [0, 1e-9]], None),     # cond 1.41e9: line 4 only
               ("inv", "f64", 0, 0, [[1e-3, 0], [0, 1e-14]], None),    # cond 1.00e14: lines 4 and 5
               ("inv", "f64", 0, 0, [[1.0, 2.0], [2.0, 4.0]], None),   # det exactly 0: no line 4
               ("solve", "f64", 0, 0, [[1.0, 2.0], [2.0, 4.0]], [1.0, 1.0])])
    assert got[0] == ("ERR", "SingularMatrixError", singular("matrix inverse", 2, 2, "1.41e9"))
    assert got[1] == ("ERR", "SingularMatrixError", singular("matrix inverse", 2, 2, "1.00e14"))
    assert got[2] == ("ERR", "SingularMatrixError", singular("matrix inverse", 2, 2))
    # solve passes inv's error through unchanged (solve.rs:66 uses `?`)
    assert got[3] == got[2]
    with pytest.raises(oracle.OracleError):
        oracle.solve(np.array([[1.0, 2.0], [2.0, 4.0]]), np.ones(2))


def _bwd(a, x, b):
    a = np.asarray(a, np.float64)
    x = np.asarray(x, np.float64).reshape(a.shape[0], -1)
    b = np.asarray(b, np.float64).reshape(a.shape[0], -1)
    r = b - a @ x
    na = np.abs(a).sum(axis=1).max()
    return max(np.abs(r[:, j]).max() / (na * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(x.shape[1]))


@pytest.mark.gpu
def test_mirror_compute_matches_oracle():
    rng = np.random.default_rng(11)
    cases, checks = [], []

    def add(case, check, want="OK"):
        cases.append(case)
        checks.append((check, want))

    for n in (3, 5, 16, 64, 300):
        a = rng.standard_normal((n, n))
        b = rng.standard_normal(n)

        def chk(g, a=a, b=b, n=n):
            x = g[1][0].ravel()
            assert _bwd(a, x, b) <= 8 * n * U64
            xo = oracle.solve(a, b)
            kappa = np.linalg.cond(a, np.inf)
            assert np.abs(x - xo).max() <= 8 * n * U64 * kappa * np.abs(xo).max()
        add(("solve", "f64", 0, 0, a, b), chk)

    a = rng.standard_normal((40, 40)).astype(np.float32)
    b = rng.standard_normal(40).astype(np.float32)
    add(("solve", "f32", 0, 0, a, b), lambda g, a=a, b=b: _bwd(a, g[1][0], b) <= 8 * 40 * 2.0 ** -24)

    a = rng.standard_normal((50, 50))
    bm = rng.standard_normal((50, 3))
    add(("solve_multiple", "f64", 0, 0, a, bm), lambda g, a=a, bm=bm: _bwd(a, g[1][0], bm) <= 8 * 50 * U64)

    for shape in ((64, 64), (7, 5), (5, 7)):
        a = rng.standard_normal(shape)

        def chk_lu(g, a=a):
            p, l, u = g[1]
            po, lo, uo = oracle.lu(a)
            assert np.array_equal(p, po)          # identical pivot sequence
            assert np.abs(l - lo).max() <= 1e-12 and np.abs(u - uo).max() <= 1e-12 * np.abs(uo).max()
            assert np.abs(p @ a - l @ u).max() <= 8 * max(a.shape) * U64 * np.abs(a).max() * max(a.shape)
        add(("lu", "f64", 0, 0, a, None), chk_lu)

    m = rng.standard_normal((30, 30))
    spd = m @ m.T + 30 * np.eye(30)
    add(("cholesky", "f64", 0, 0, spd, None),
        lambda g, spd=spd: np.abs(g[1][0] - oracle.cholesky(spd)).max() <= 1e-12 * np.sqrt(np.abs(spd).max()))

    for lower, unit in ((1, 0), (0, 0), (1, 1), (0, 1)):
        t = rng.standard_normal((50, 50)) + 8 * np.eye(50)
        t = np.tril(t) if lower else np.triu(t)
        b = rng.standard_normal(50)

        def chk_t(g, t=t, b=b, lower=lower, unit=unit):
            xo = oracle.solve_triangular(t, b, bool(lower), bool(unit))
            assert np.abs(g[1][0].ravel() - xo).max() <= 1e-12 * np.abs(xo).max()
        add(("solve_triangular", "f64", lower, unit, t, b), chk_t)
        # f32: the mirror hands f32 to scirs_gpu_strsv (no upcast); the oracle's
        # f32 loop within 8 n 2^-24 kappa
        t32, b32 = t.astype(np.float32), b.astype(np.float32)

        def chk_t32(g, t32=t32, b32=b32, lower=lower, unit=unit):
            xo = oracle.solve_triangular(t32, b32, bool(lower), bool(unit))
            te = t32.astype(np.float64)
            if unit:
                np.fill_diagonal(te, 1.0)
            bound = 8 * 50 * 2.0 ** -24 * np.linalg.cond(te, np.inf)
            assert np.abs(g[1][0].ravel() - xo).max() <= bound * np.abs(xo).max()
        add(("solve_triangular", "f32", lower, unit, t32, b32), chk_t32)

    a6 = rng.standard_normal((6, 6))
    add(("det", "f64", 0, 0, a6, None),
        lambda g, a6=a6: abs(g[1][0][0, 0] - np.linalg.det(a6)) <= 1e-12 * abs(np.linalg.det(a6)))
    add(("det", "f64", 0, 0, np.diag([1.0, 2, 3, 0, 5]), None), lambda g: g[1][0][0, 0] == 0.0)
    for n in (3, 10):
        a = rng.standard_normal((n, n))
        add(("inv", "f64", 0, 0, a, None),
            lambda g, a=a, n=n: np.abs(g[1][0] @ a - np.eye(n)).max() <= 1e-12 * np.linalg.cond(a))

    # breakdowns: same error variants and texts as the reference
    sing = rng.standard_normal((8, 8))
    sing[:, 3] = 0.0
    add(("solve", "f64", 0, 0, sing, np.ones(8)),
        lambda g: g == ("ERR", "SingularMatrixError", singular("linear system solve", 8, 8)), "ERR")
    add(("solve_multiple", "f64", 0, 0, sing, np.ones((8, 2))),
        lambda g: g == ("ERR", "SingularMatrixError", singular("multiple linear systems solve", 8, 8)), "ERR")
    add(("lu", "f64", 0, 0, sing, None),
        lambda g: g == ("ERR", "SingularMatrixError", singular("LU decomposition", 8, 8)), "ERR")
    add(("inv", "f64", 0, 0, sing, None),
        lambda g: g == ("ERR", "SingularMatrixError", singular("matrix inverse via solve", 8, 8)), "ERR")
    npd = spd.copy()
    npd[10, 10] = -1.0
    add(("cholesky", "f64", 0, 0, npd, None),
        lambda g: g == ("ERR", "NonPositiveDefiniteError", not_pd("Cholesky decomposition", 30, 30)), "ERR")

    got = run(cases)
    for case, g, (chk, want) in zip(cases, got, checks):
        label = (case[0], case[1], np.shape(case[4]))
        assert g[0] == want, (label, g)
        res = chk(g)
        assert res is None or res, (label, g)


def test_mixed_precision_dropins_validation_texts():
    """mixed_precision_solve / iterative_refinement_solve (mixed_precision/
    adaptive.rs:49-151, 282-378): the reference's own shape-error texts (the
    `{:?}` of the shape slice) before any device work."""
    a34 = np.ones((3, 4))
    a33 = np.eye(3)
    got = run([
        ("mixed_precision_solve", "f64", 0, 0, a34, np.ones(3)),
        ("mixed_precision_solve", "f64", 0, 0, a33, np.ones(2)),
        ("iterative_refinement_solve", "f64", 0, 0, a34, np.ones(3)),
        ("iterative_refinement_solve", "f64", 0, 0, a33, np.ones(2)),
    ])
    assert got[0] == ("ERR", "ShapeError", "Shape error: Matrix must be square, got shape [3, 4]")
    assert got[1] == ("ERR", "ShapeError", "Shape error: Matrix rows (3) must match vector length (2)")
    assert got[2] == ("ERR", "ShapeError",
                      "Shape error: Expected square matrix for iterative refinement, got [3, 4]")
    assert got[3] == ("ERR", "ShapeError",
                      "Shape error: Matrix shape [3, 3] is incompatible with vector length 2")


@pytest.mark.gpu
def test_mixed_precision_dropins_match_oracle():
    """The drop-ins on the GPU: the FP64 elimination of mixed_precision_solve and
    the FP32-factor + FP64-refinement form of iterative_refinement_solve both give
    the FP64 solution (backward error <= 8 n 2^-53, the oracle's solution within
    8 n u kappa); an FP32 elimination gives the FP32 one; a singular system fails
    as the reference's elimination (`Matrix is singular`) and working-precision
    solve (`linear system solve`) do."""
    rng = np.random.default_rng(77)
    n = 200
    a = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    b = rng.standard_normal(n)
    xo = oracle.solve(a, b)
    kappa = np.linalg.cond(a, np.inf)
    sing = rng.standard_normal((8, 8))
    sing[:, 5] = 0.0
    a3 = rng.standard_normal((3, 3)) + 3 * np.eye(3)
    b3 = rng.standard_normal(3)
    got = run([
        ("mixed_precision_solve", "f64", 0, 0, a, b),
        ("mixed_precision_solve", "f64", 1, 0, a, b),
        ("iterative_refinement_solve", "f64", 0, 0, a, b),
        ("iterative_refinement_solve", "f64", 1, 0, a, b),
        ("iterative_refinement_solve", "f64", 0, 3, a3, b3),
        ("mixed_precision_solve", "f64", 0, 0, sing, np.ones(8)),
        ("iterative_refinement_solve", "f64", 0, 0, sing, np.ones(8)),
    ])
    for i in (0, 2, 3):
        assert got[i][0] == "OK", got[i]
        x = got[i][1][0].ravel()
        assert _bwd(a, x, b) <= 8 * n * U64, i
        assert np.abs(x - xo).max() <= 8 * n * U64 * kappa * np.abs(xo).max(), i
    x32 = got[1][1][0].ravel()
    assert np.abs(x32 - xo).max() <= 8 * n * 2.0 ** -24 * kappa * np.abs(xo).max()
    assert got[4][0] == "OK"
    # n <= 3: FP32 inverse-branch solve, then the reference's FP64 refinement loop
    assert np.abs(got[4][1][0].ravel() - np.linalg.solve(a3, b3)).max() <= 1e-12
    assert got[5] == ("ERR", "SingularMatrixError", "Singular matrix error: Matrix is singular")
    assert got[6] == ("ERR", "SingularMatrixError", singular("linear system solve", 8, 8))


def test_compat_solve_options_and_texts():
    """compat_solve (compat.rs:461-499): the finiteness scan and the options the
    reference does not support keep its errors (no device work)."""
    a = np.eye(3)
    b = np.ones((3, 1))
    bad = a.copy()
    bad[1, 2] = np.nan
    got = run([
        ("compat_solve", "f64", 4, 0, bad, b),         # check_finite
        ("compat_solve", "f64", 1, 0, a, b),           # lower
        ("compat_solve", "f64", 2, 0, a, b),           # transposed
        ("compat_solve", "f64", 0, 2, a, b),           # assume_a = "sym"
        ("compat_solve", "f64", 1, 1, a, b),           # "pos" with lower
    ])
    assert got[0] == ("ERR", "ValueError", "Value error: Matrix contains non-finite values")
    for g in got[1:]:
        assert g == ("ERR", "NotImplementedError", "Not implemented: Specialized solve options not yet supported")


@pytest.mark.gpu
def test_compat_solve_general_and_pos():
    """compat_solve with no options is solve_multiple; assume_a = "pos" runs the
    GPU Cholesky solve (LAPACK's solution within 8 n u kappa) and reports a
    non-positive-definite matrix as the Cholesky decomposition does."""
    rng = np.random.default_rng(91)
    n, nrhs = 150, 3
    g = rng.standard_normal((n, n))
    s = g @ g.T / n + np.eye(n)
    a = rng.standard_normal((n, n))
    b = rng.standard_normal((n, nrhs))
    npd = s.copy()
    npd[7, 7] = -1.0
    got = run([
        ("compat_solve", "f64", 4, 0, a, b),
        ("compat_solve", "f64", 4, 1, s, b),
        ("compat_solve", "f32", 0, 1, s, b),
        ("compat_solve", "f64", 0, 1, npd, b),
    ])
    for i, m in ((0, a), (1, s)):
        assert got[i][0] == "OK", got[i]
        x = got[i][1][0]
        xo = np.linalg.solve(m, b)
        kappa = np.linalg.cond(m, np.inf)
        assert np.abs(x - xo).max() <= 8 * n * U64 * kappa * np.abs(xo).max(), i
    x32 = got[2][1][0]
    s32, b32 = s.astype(np.float32), b.astype(np.float32)
    xo32 = oracle.cholesky_solve(s32, b32)   # the composition in f32
    kappa = np.linalg.cond(s, np.inf)
    assert np.abs(x32 - xo32).max() <= 8 * n * 2.0 ** -24 * kappa * np.abs(xo32).max()
    assert got[3] == ("ERR", "NonPositiveDefiniteError", not_pd("Cholesky decomposition", n, n))


def test_qr_lstsq_validation_texts():
    """qr (decomposition.rs:208-249, lapack.rs:185-190) and lstsq
    (solve.rs:234-250, validation.rs:247-261): errors raised before any device
    work keep the reference's kinds and texts."""
    bad = np.ones((3, 2))
    bad[2, 1] = np.nan
    got = run([
        ("qr", "f64", 0, 0, np.zeros((0, 0)), None),
        ("qr", "f64", 0, 0, bad, None),
        ("qr", "f64", 0, 0, np.ones((2, 3)), None),
        ("lstsq", "f64", 0, 0, bad, np.ones(3)),
        ("lstsq", "f64", 0, 0, np.ones((3, 2)), np.array([1.0, INF, 2.0])),
        ("lstsq", "f64", 0, 0, np.ones((2, 3)), np.ones(2)),
    ])
    assert got[0] == ("ERR", "ShapeError", "Shape error: QR decomposition failed: Input matrix cannot be empty")
    assert got[1] == ("ERR", "InvalidInputError",
                      "Invalid input error: QR decomposition failed: Matrix contains non-finite values")
    assert got[2] == ("ERR", "ComputationError", "Computation error: QR decomposition requires rows >= columns")
    assert got[3] == ("ERR", "InvalidInputError",
                      "Invalid input error: Least squares solve failed: Matrix contains non-finite values")
    assert got[4] == ("ERR", "InvalidInputError",
                      "Invalid input error: Least squares solve failed: Vector contains non-finite values")
    assert got[5][0] == "ERR" and got[5][1] == "NotImplementedError"


@pytest.mark.gpu
def test_qr_lstsq_results():
    """qr and lstsq through the mirror: the oracle's Q, R (lapack.rs:171-253) and
    x, residuals (solve.rs:234-290) within 64 rows u; rank = cols, s empty; a
    rank-deficient R reports the back substitution's singular error."""
    from oracle import oracle
    rng = np.random.default_rng(17)
    a = rng.standard_normal((60, 24))
    b = rng.standard_normal(60)
    dep = np.array([[1.0, 0.0, 2.0], [3.0, 0.0, 1.0], [4.0, 0.0, 5.0], [2.0, 0.0, 7.0]])
    got = run([
        ("qr", "f64", 0, 0, a, None),
        ("lstsq", "f64", 0, 0, a, b),
        ("lstsq", "f64", 0, 0, [[1.0, 1.0], [1.0, 2.0], [1.0, 3.0]], [6.0, 9.0, 12.0]),
        ("lstsq", "f64", 0, 0, dep, np.ones(4)),
    ])
    tol = 64 * 60 * U64 * np.abs(a).max()
    qo, ro = oracle.qr_factor(a)
    assert got[0][0] == "OK"
    assert np.abs(got[0][1][0] - qo).max() <= tol and np.abs(got[0][1][1] - ro).max() <= tol
    xo, reso = oracle.lstsq(a, b)
    x, res, rank, s = got[1][1]
    assert np.abs(x[:, 0] - xo).max() <= 1e-9 * np.abs(xo).max()
    assert abs(res[0, 0] - reso) <= 1e-9 * reso
    assert rank[0, 0] == 24 and s.size == 0
    assert np.allclose(got[2][1][0][:, 0], [3.0, 3.0], atol=1e-12)
    assert got[3] == ("ERR", "SingularMatrixError",
                      singular("triangular system solve (back substitution)", 3, 3, "1.00e16"))


def test_batch_validation_texts():
    """batch/mod.rs:70-75, 147-152, 230-262: shape and axis errors keep the
    reference's kinds and texts (no device work)."""
    a = np.ones((4, 3))   # two 2 x 3 matrices
    got = run([
        ("batch_matmul", "f64", 2, 0, a, np.ones((2, 2))),
        ("batch_matvec", "f64", 2, 0, a, np.ones(2)),
        ("batch_add", "f64", 2, 0, a, np.ones(3)),
        ("batch_add", "f64", 2, 1, a, np.ones(2)),
        ("batch_add", "f64", 2, 2, a, np.ones(3)),
    ])
    assert got[0] == ("ERR", "ShapeError", "Shape error: Inner dimensions mismatch for batch_matmul: 3 vs 2")
    assert got[1] == ("ERR", "ShapeError",
                      "Shape error: Dimension mismatch for batch_matvec: matrix width 3 does not match vector length 2")
    assert got[2] == ("ERR", "ShapeError",
                      "Shape error: Column-wise batch_add requires vector length 2 to match number of rows 2, got 3")
    assert got[3] == ("ERR", "ShapeError",
                      "Shape error: Row-wise batch_add requires vector length 3 to match number of columns 3, got 2")
    assert got[4] == ("ERR", "InvalidInputError",
                      "Invalid input error: Invalid axis 2: must be 0 (column-wise) or 1 (row-wise)")


@pytest.mark.gpu
def test_batch_results_bit_identical():
    """The mirror's batch operations return the oracle's bits (batch/mod.rs:62-336)."""
    from oracle import oracle
    rng = np.random.default_rng(23)
    a = rng.standard_normal((3, 17, 29))
    b = rng.standard_normal((29, 13))
    x = rng.standard_normal(29)
    vr = rng.standard_normal(17)
    flat = a.reshape(3 * 17, 29)
    got = run([
        ("batch_matmul", "f64", 3, 0, flat, b),
        ("batch_matvec", "f64", 3, 0, flat, x),
        ("batch_add", "f64", 3, 0, flat, vr),
        ("batch_add", "f64", 3, 1, flat, x),
        ("batch_sum", "f64", 3, 0, flat, None),
    ])
    assert all(g[0] == "OK" for g in got), got
    assert np.array_equal(got[0][1][0], oracle.batch_matmul(a, b).reshape(51, 13))
    assert np.array_equal(got[1][1][0], oracle.batch_matvec(a, x))
    assert np.array_equal(got[2][1][0], oracle.batch_add(a, vr, 0).reshape(51, 29))
    assert np.array_equal(got[3][1][0], oracle.batch_add(a, x, 1).reshape(51, 29))
    assert np.array_equal(got[4][1][0], oracle.batch_sum(a))


def test_gmres_validation_texts():
    """gmres (solvers/iterative.rs:327-350): shape errors keep the reference's texts."""
    got = run([
        ("gmres", "f64", 0, 0, np.ones((2, 3)), np.ones(2)),
        ("gmres", "f64", 0, 0, np.eye(3), np.ones(2)),
    ])
    assert got[0] == ("ERR", "ShapeError", "Shape error: Matrix A must be square")
    assert got[1] == ("ERR", "ShapeError", "Shape error: Vector b must have same length as matrix dimension")


def test_gmres_restart_zero_is_an_error():
    """restart = Some(0): the reference divides by restart.min(n) = 0 and panics
    (solvers/iterative.rs:361-366); the mirror returns an error, before any
    device work."""
    got = run([("gmres", "f64", 0, -1, np.eye(3), np.ones(3))])
    assert got[0] == ("ERR", "InvalidInputError", "Invalid input error: GMRES restart must be positive")


@pytest.mark.gpu
def test_gmres_results():
    """gmres through the mirror: the oracle's iteration count, flag and x."""
    from oracle import oracle
    rng = np.random.default_rng(31)
    n = 80
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 2.0 * np.eye(n)
    b = rng.standard_normal(n)
    got = run([("gmres", "f64", 0, 0, a, b), ("gmres", "f64", 7, 0, a, b), ("gmres", "f64", 0, 0,
                                                                              [[1.0, 2.0], [3.0, 4.0]], [5.0, 11.0])])
    for g, kw in ((got[0], {}), (got[1], {"max_iter": 7})):
        assert g[0] == "OK", g
        x, it, res, conv = g[1]
        xo, ito, reso, convo = oracle.gmres(a, b, **kw)
        assert int(it[0, 0]) == ito and bool(conv[0, 0]) == convo
        assert np.abs(x[:, 0] - xo).max() <= 1e-9 * max(1.0, np.abs(xo).max())
    assert np.allclose(got[2][1][0][:, 0], [1.0, 2.0], atol=1e-10)


def test_preconditioned_gmres_validation_texts():
    """preconditioned_gmres (preconditioners.rs:947-952): the reference's shape
    text; DiagonalPreconditioner's own shape check (preconditioners.rs:211-216);
    restart 0 (a division by zero in the reference) is an error.  No device
    work is reached."""
    got = run([
        ("preconditioned_gmres", "f64", 1, 0, np.ones((2, 3)), np.ones(2)),
        ("preconditioned_gmres", "f64", 1, 0, np.eye(3), np.ones(2)),
        ("preconditioned_gmres", "f64", 0, 0, np.ones((2, 3)), np.ones(2)),
        ("preconditioned_gmres", "f64", 1, -1, np.eye(3), np.ones(3)),
    ])
    assert got[0] == ("ERR", "ShapeError", "Shape error: Matrix and RHS dimensions must be compatible")
    assert got[1] == ("ERR", "ShapeError", "Shape error: Matrix and RHS dimensions must be compatible")
    assert got[2] == ("ERR", "ShapeError", "Shape error: Diagonal preconditioner requires square matrix")
    assert got[3] == ("ERR", "InvalidInputError", "Invalid input error: GMRES restart must be positive")


@pytest.mark.gpu
def test_preconditioned_gmres_results():
    """Through the mirror, both preconditioners solve A x = b (the reference's
    stub would not): Jacobi gives the oracle's right-preconditioned iterate,
    the LU preconditioner LAPACK's solution."""
    from oracle import oracle
    rng = np.random.default_rng(41)
    n = 90
    d = np.logspace(-2, 2, n)
    a = d[:, None] * (np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n))
    b = rng.standard_normal(n)
    got = run([("preconditioned_gmres", "f64", 0, 10, a, b), ("preconditioned_gmres", "f64", 1, 10, a, b)])
    xo, ito, _, convo = oracle.pgmres(a, b, 1, max_iter=1000, tol=1e-10, restart=10)
    assert convo and got[0][0] == "OK", got[0]
    kappa = np.linalg.cond(a)
    assert np.abs(got[0][1][0][:, 0] - xo).max() <= 1e-10 * kappa * max(1.0, np.abs(xo).max())
    xs = np.linalg.solve(a, b)
    assert got[1][0] == "OK", got[1]
    assert np.abs(got[1][1][0][:, 0] - xs).max() <= 1e-10 * kappa * max(1.0, np.abs(xs).max())
