"""Writes the golden fixtures under tests/golden/.

Two kinds of data, both inputs plus expected outputs:

1. ``reference_kats.json`` — the known-answer cases that scirs2-linalg's own
   tests hold for the dense-solve path (values copied as data, each with the
   file:line of the test that asserts them).
2. ``lapack_n{5,8,16,64,512}.npz`` — seeded N(0,1) systems (LU) and
   SPD systems A = G G^T / n + I (Cholesky), solved by LAPACK through scipy
   1.15 (dgetrf/dgetrs, dpotrf/dpotrs).  LAPACK's idamax picks the first
   maximum, the same tie rule as lapack.rs:95-104, so for these inputs the
   pivot sequence must equal the reference's, and the factors agree to
   rounding.

Run:  python tests/golden/gen_golden.py
"""
import json
import os

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))

KATS = {
    "lu_factor_2x2": {
        "source": "scirs2-linalg/src/lapack.rs:688-734",
        "a": [[2.0, 1.0], [4.0, 3.0]],
        "lu": [[4.0, 3.0], [0.5, -0.5]],
        "piv": [1, 0],
    },
    "cholesky_2x2": {
        "source": "scirs2-linalg/src/lapack.rs:737-754; decomposition.rs:757-776",
        "a": [[4.0, 2.0], [2.0, 5.0]],
        "l": [[2.0, 0.0], [1.0, 2.0]],
    },
    "cholesky_3x3_textbook": {
        "source": "scirs2-linalg/tests/scipy_validation.rs:131-152",
        "a": [[4.0, 12.0, -16.0], [12.0, 37.0, -43.0], [-16.0, -43.0, 98.0]],
        "l": [[2.0, 0.0, 0.0], [6.0, 1.0, 0.0], [-8.0, 5.0, 3.0]],
    },
    "cholesky_spd_set": {
        "source": "scirs2-linalg/tests/scipy_compat_comprehensive_tests.rs:359-390",
        "mats": [
            [[4.0, 2.0], [2.0, 3.0]],
            [[2.0, 1.0, 0.0], [1.0, 2.0, 1.0], [0.0, 1.0, 2.0]],
            [[5.0, 2.0, 1.0], [2.0, 4.0, 2.0], [1.0, 2.0, 3.0]],
        ],
        "tol": 1e-8,
    },
    "lu_property_set": {
        "source": "scirs2-linalg/tests/scipy_validation.rs:49-75; "
                  "scipy_compat_comprehensive_tests.rs:246-280; scipy_compat_validation.rs:365-374",
        "mats": [
            [[2.0, 1.0, 1.0], [4.0, -6.0, 0.0], [-2.0, 7.0, 2.0]],
            [[2.0, 1.0], [1.0, 1.0]],
            [[4.0, 2.0, 1.0], [2.0, 5.0, 3.0], [1.0, 3.0, 6.0]],
            [[1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 10.0]],
        ],
        "tol": 1e-10,
    },
    "solve_cases": {
        "source": "scirs2-linalg/src/solve.rs:458-472; tests/scipy_validation.rs:178-189; "
                  "src/mixed_precision/adaptive.rs:270-281; src/solvers/iterative.rs:688-700",
        "cases": [
            {"a": [[1.0, 0.0], [0.0, 1.0]], "b": [2.0, 3.0], "x": [2.0, 3.0]},
            {"a": [[1.0, 2.0], [3.0, 4.0]], "b": [5.0, 11.0], "x": [1.0, 2.0]},
            {"a": [[3.0, 1.0], [1.0, 2.0]], "b": [9.0, 8.0], "x": [2.0, 3.0]},
        ],
        "tol": 1e-10,
    },
    "solve_triangular_cases": {
        "source": "scirs2-linalg/src/solve.rs:474-507",
        "cases": [
            {"a": [[1.0, 0.0], [2.0, 3.0]], "b": [2.0, 8.0], "lower": True, "unit": False,
             "x": [2.0, 4.0 / 3.0]},
            {"a": [[1.0, 0.0], [2.0, 1.0]], "b": [2.0, 6.0], "lower": True, "unit": True,
             "x": [2.0, 2.0]},
            {"a": [[3.0, 2.0], [0.0, 1.0]], "b": [8.0, 2.0], "lower": False, "unit": False,
             "x": [4.0 / 3.0, 2.0]},
            {"a": [[1.0, 2.0], [0.0, 1.0]], "b": [6.0, 2.0], "lower": False, "unit": True,
             "x": [2.0, 2.0]},
        ],
    },
    "solve_residual_cases": {
        "source": "scirs2-linalg/tests/scipy_compat_comprehensive_tests.rs:447-495; "
                  "tests/scipy_compat_validation.rs:455-476",
        "cases": [
            {"a": [[3.0, 1.0], [1.0, 2.0]], "b": [[4.0], [3.0]], "tol": 1e-8},
            {"a": [[3.0, 1.0], [1.0, 2.0]], "b": [[4.0, 1.0], [3.0, 2.0]], "tol": 1e-8},
            {"a": [[4.0, 2.0, 1.0], [2.0, 5.0, 3.0], [1.0, 3.0, 6.0]], "b": [[1.0], [2.0], [3.0]],
             "tol": 1e-8},
            {"a": [[10.0, 1.0, 1.0], [1.0, 10.0, 1.0], [1.0, 1.0, 10.0]], "b": [[1.0], [1.0], [1.0]],
             "tol": 1e-12},
        ],
    },
    "singular_solve": {
        "source": "scirs2-linalg/tests/error_diagnostics_tests.rs:23-41",
        "a": [[1.0, 2.0], [2.0, 4.0]],
        "b": [1.0, 2.0],
        "message_contains": ["Singular matrix", "Matrix shape", "Consider the following"],
    },
    "mixed_precision_2x2": {
        "source": "scirs2-linalg/src/mixed_precision/adaptive.rs:595-606",
        "a": [[2.0, 1.0], [1.0, 3.0]],
        "b": [5.0, 8.0],
        "x": [1.4, 2.2],
        "tol": 1e-4,
    },
}


def write_kats():
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(KATS, f, indent=1)


def write_lapack(n: int, seed: int):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    bm = rng.standard_normal((n, 4))
    lu, ipiv = sla.lu_factor(a)          # dgetrf, LAPACK ipiv (0-based swap list)
    perm = np.arange(n)
    for k, p in enumerate(ipiv):        # swap list -> reference-style permutation
        perm[k], perm[p] = perm[p], perm[k]
    x = sla.lu_solve((lu, ipiv), b)
    xm = sla.lu_solve((lu, ipiv), bm)
    g = rng.standard_normal((n, n))
    spd = g @ g.T / n + np.eye(n)
    spd = 0.5 * (spd + spd.T)
    lchol = sla.cholesky(spd, lower=True)   # dpotrf
    xs = sla.cho_solve((lchol, True), b)    # dpotrs
    out = dict(seed=seed, a=a, b=b, bm=bm, perm=perm, x=x, xm=xm)
    if n <= 64:  # full factors and the SPD case only at small n (keeps the repo small)
        out.update(lu=lu, spd=spd, lchol=lchol, xs=xs)
    np.savez_compressed(os.path.join(HERE, f"lapack_n{n}.npz"), **out)


def main():
    write_kats()
    for n, seed in ((5, 11), (8, 12), (16, 13), (64, 14), (512, 15)):
        write_lapack(n, seed)


if __name__ == "__main__":
    main()
