"""ORACLE — test infrastructure only.

ctypes front end of ``oracle/_build/liboracle.so`` (built from
``oracle/ref_linalg.c`` by ``oracle/Makefile``), the plain-C restatement of
scirs2-linalg's CPU solve path:

* ``lu_factor``        -> scirs2-linalg/src/lapack.rs:72-147
* ``cholesky``         -> scirs2-linalg/src/lapack.rs:632-679
* ``solve_triangular`` -> scirs2-linalg/src/solve.rs:139-202
* ``solve``            -> scirs2-linalg/src/solve.rs:54-111 (+ basic.rs:117-186)
* ``solve_multiple``   -> scirs2-linalg/src/solve.rs:366-424
* ``lu``               -> scirs2-linalg/src/decomposition.rs:157-180 (P, L, U build)
* ``qr_factor``        -> scirs2-linalg/src/lapack.rs:171-253 (f64)
* ``lstsq``            -> scirs2-linalg/src/solve.rs:234-290, rows >= cols (f64)
* ``batch_matmul``, ``batch_matvec``, ``batch_add``, ``batch_sum``
                       -> scirs2-linalg/src/batch/mod.rs:62-336
* ``gmres``            -> scirs2-linalg/src/solvers/iterative.rs:317-450 (f64)

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module; it is the checker, never the product.
Pinning of each function is listed in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


class OracleError(RuntimeError):
    """Raised when the oracle reports a singular / non-PD / empty input."""

    def __init__(self, kind: str, code: int):
        super().__init__(f"{kind} (code {code})")
        self.kind = kind
        self.code = code


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    L = ctypes.c_long
    for s in ("d", "s"):
        getattr(lib, f"oracle_lu_factor_{s}").argtypes = [I, I, P, L, P, P, P, P]
        getattr(lib, f"oracle_cholesky_{s}").argtypes = [I, P, L, P]
        getattr(lib, f"oracle_solve_triangular_{s}").argtypes = [I, P, L, P, P, I, I]
        getattr(lib, f"oracle_solve_{s}").argtypes = [I, P, L, P, P]
        getattr(lib, f"oracle_solve_multiple_{s}").argtypes = [I, I, P, L, P, L, P, L]
    for s in ("d", "s"):
        getattr(lib, f"oracle_batch_matmul_{s}").argtypes = [I, I, I, I, P, P, P]
        getattr(lib, f"oracle_batch_matvec_{s}").argtypes = [I, I, I, P, P, P]
        getattr(lib, f"oracle_batch_add_{s}").argtypes = [I, I, I, P, P, I, P]
        getattr(lib, f"oracle_batch_sum_{s}").argtypes = [I, I, I, P, P]
    lib.oracle_gmres_d.argtypes = [I, P, L, P, P, I, ctypes.c_double, I, P, P, P, P]
    lib.oracle_pgmres_d.argtypes = [I, P, L, P, P, I, ctypes.c_double, I, I, P, P, P, P, P, P]
    lib.oracle_qr_factor_d.argtypes = [I, I, P, L, P, P]
    lib.oracle_lstsq_d.argtypes = [I, I, P, L, P, P, P]
    _lib = lib
    return lib


def _sfx(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "d"
    if dt == np.float32:
        return "s"
    raise TypeError(f"oracle supports float32/float64, got {dt}")


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lu_factor(a):
    """Packed LU, permutation vector and sign, exactly as lapack.rs:72-147.

    Returns (lu, piv, sign).  Raises OracleError('singular', k) with the
    1-based column k where the reference's |pivot| < eps test fires.
    """
    a = np.asarray(a)
    dt = a.dtype if a.dtype in (np.float32, np.float64) else np.float64
    a = _c(a, dt)
    n, m = a.shape
    lu = np.empty_like(a)
    piv = np.empty(n, dtype=np.int64)
    sign = np.zeros(1, dtype=dt)
    cond = np.zeros(1, dtype=np.float64)
    rc = getattr(_load(), f"oracle_lu_factor_{_sfx(dt)}")(
        n, m, _ptr(a), m, _ptr(lu), _ptr(piv), _ptr(sign), _ptr(cond))
    if rc > 0:
        err = OracleError("singular", rc)
        err.cond_estimate = None if cond[0] < 0 else float(cond[0])
        raise err
    if rc < 0:
        raise OracleError("empty", rc)
    return lu, piv, dt.type(sign[0]) if hasattr(dt, "type") else sign[0]


def lu(a):
    """(P, L, U) built from lu_factor exactly as decomposition.rs:157-180."""
    a = np.asarray(a)
    packed, piv, _ = lu_factor(a)
    n, m = packed.shape
    k = min(n, m)
    p = np.zeros((n, n), dtype=packed.dtype)
    p[np.arange(n), piv] = 1
    l = np.zeros((n, k), dtype=packed.dtype)
    for i in range(n):
        l[i, : min(i, m)] = packed[i, : min(i, m)]
        if i < m:
            l[i, i] = 1
    u = np.zeros((k, m), dtype=packed.dtype)
    for i in range(k):
        u[i, i:] = packed[i, i:]
    return p, l, u


def cholesky(a):
    """Lower Cholesky factor (lapack.rs:632-679); reads only the lower triangle."""
    a = np.asarray(a)
    dt = a.dtype if a.dtype in (np.float32, np.float64) else np.float64
    a = _c(a, dt)
    n = a.shape[0]
    l = np.empty_like(a)
    rc = getattr(_load(), f"oracle_cholesky_{_sfx(dt)}")(n, _ptr(a), n, _ptr(l))
    if rc > 0:
        raise OracleError("not_positive_definite", rc)
    if rc < 0:
        raise OracleError("empty", rc)
    return l


def solve_triangular(a, b, lower: bool, unit_diagonal: bool):
    a = np.asarray(a)
    dt = a.dtype if a.dtype in (np.float32, np.float64) else np.float64
    a = _c(a, dt)
    b = _c(b, dt)
    n = a.shape[0]
    x = np.empty(n, dtype=dt)
    rc = getattr(_load(), f"oracle_solve_triangular_{_sfx(dt)}")(
        n, _ptr(a), n, _ptr(b), _ptr(x), int(lower), int(unit_diagonal))
    if rc > 0:
        raise OracleError("singular", rc)
    return x


def solve(a, b):
    a = np.asarray(a)
    dt = a.dtype if a.dtype in (np.float32, np.float64) else np.float64
    a = _c(a, dt)
    b = _c(b, dt)
    n = a.shape[0]
    x = np.empty(n, dtype=dt)
    rc = getattr(_load(), f"oracle_solve_{_sfx(dt)}")(n, _ptr(a), n, _ptr(b), _ptr(x))
    if rc > 0:
        raise OracleError("singular", rc)
    if rc < 0:
        raise OracleError("empty", rc)
    return x


def solve_multiple(a, b):
    a = np.asarray(a)
    dt = a.dtype if a.dtype in (np.float32, np.float64) else np.float64
    a = _c(a, dt)
    b = _c(b, dt)
    n = a.shape[0]
    nrhs = b.shape[1]
    x = np.empty((n, nrhs), dtype=dt)
    rc = getattr(_load(), f"oracle_solve_multiple_{_sfx(dt)}")(
        n, nrhs, _ptr(a), n, _ptr(b), nrhs, _ptr(x), nrhs)
    if rc > 0:
        raise OracleError("singular", rc)
    if rc < 0:
        raise OracleError("empty", rc)
    return x


def cholesky_solve(a, b):
    """Composition used as the Cholesky-solve oracle: cholesky (lapack.rs:632)
    then solve_triangular(L, b, lower) and solve_triangular(L^T, y, upper)
    (solve.rs:139-202), column by column for a 2-D b."""
    l = cholesky(a)
    b = np.asarray(b, dtype=l.dtype)
    cols = b.reshape(b.shape[0], -1)
    out = np.empty_like(cols)
    lt = np.ascontiguousarray(l.T)
    for j in range(cols.shape[1]):
        y = solve_triangular(l, cols[:, j], True, False)
        out[:, j] = solve_triangular(lt, y, False, False)
    return out.reshape(b.shape)


def qr_factor(a):
    """(Q, R) of an f64 rows x cols matrix (rows >= cols) as lapack.rs:171-253."""
    a = _c(a, np.float64)
    rows, cols = a.shape
    q = np.empty((rows, rows))
    r = np.empty((rows, cols))
    rc = _load().oracle_qr_factor_d(rows, cols, _ptr(a), cols, _ptr(q), _ptr(r))
    if rc:
        raise OracleError("empty" if rc == -1 else "shape", rc)
    return q, r


def lstsq(a, b):
    """(x, residuals) of the rows >= cols branch of solve.rs:234-290 (f64).

    Raises OracleError('singular', k) where the back substitution on R hits
    |r_kk| < eps (1-based row k)."""
    a = _c(a, np.float64)
    b = _c(b, np.float64)
    rows, cols = a.shape
    x = np.empty(cols)
    res = np.zeros(1)
    rc = _load().oracle_lstsq_d(rows, cols, _ptr(a), cols, _ptr(b), _ptr(x), _ptr(res))
    if rc > 0:
        raise OracleError("singular", rc)
    if rc < 0:
        raise OracleError("empty" if rc == -1 else "shape", rc)
    return x, float(res[0])


def _bt(a):
    a = np.asarray(a)
    dt = np.float32 if a.dtype == np.float32 else np.float64
    return _c(a, dt), dt


def batch_matmul(a, b):
    """batch/mod.rs:62-107: (batch, m, k) x (k, n) -> (batch, m, n), sums in k order."""
    a, dt = _bt(a)
    b = _c(b, dt)
    p, m, k = a.shape
    n = b.shape[1]
    c = np.empty((p, m, n), dtype=dt)
    getattr(_load(), f"oracle_batch_matmul_{_sfx(dt)}")(p, m, k, n, _ptr(a), _ptr(b), _ptr(c))
    return c


def batch_matvec(a, x):
    """batch/mod.rs:139-172: (batch, m, n) x (n,) -> (batch, m)."""
    a, dt = _bt(a)
    x = _c(x, dt)
    p, m, n = a.shape
    y = np.empty((p, m), dtype=dt)
    getattr(_load(), f"oracle_batch_matvec_{_sfx(dt)}")(p, m, n, _ptr(a), _ptr(x), _ptr(y))
    return y


def batch_add(a, v, axis):
    """batch/mod.rs:220-291: v broadcast along axis 0 (rows) or 1 (columns)."""
    a, dt = _bt(a)
    v = _c(v, dt)
    p, m, n = a.shape
    out = np.empty_like(a)
    getattr(_load(), f"oracle_batch_add_{_sfx(dt)}")(p, m, n, _ptr(a), _ptr(v), int(axis), _ptr(out))
    return out


def batch_sum(a):
    """batch/mod.rs:326-336: the sum over the batch axis, added in batch order."""
    a, dt = _bt(a)
    p, m, n = a.shape
    out = np.empty((m, n), dtype=dt)
    getattr(_load(), f"oracle_batch_sum_{_sfx(dt)}")(p, m, n, _ptr(a), _ptr(out))
    return out


def gmres(a, b, x0=None, max_iter=1000, tol=1e-10, restart=0):
    """solvers/iterative.rs:317-450 (f64): (x, iterations, residual_norm, converged).

    Raises OracleError('singular', k) where the least-squares solve of the
    Hessenberg system meets |r_kk| < eps."""
    a = _c(a, np.float64)
    b = _c(b, np.float64)
    n = a.shape[0]
    x = np.empty(n)
    it = np.zeros(1, dtype=np.int32)
    conv = np.zeros(1, dtype=np.int32)
    res = np.zeros(1)
    x0p = None if x0 is None else _c(x0, np.float64)
    rc = _load().oracle_gmres_d(n, _ptr(a), n, _ptr(b), None if x0p is None else _ptr(x0p), int(max_iter),
                                float(tol), int(restart), _ptr(x), _ptr(it), _ptr(res), _ptr(conv))
    if rc > 0:
        raise OracleError("singular", rc)
    if rc < 0:
        raise OracleError("memory", rc)
    return x, int(it[0]), float(res[0]), bool(conv[0])


def pgmres(a, b, kind, lu=None, perm=None, x0=None, max_iter=1000, tol=1e-10, restart=0):
    """Right-preconditioned restatement of solvers/iterative.rs:317-450 (the
    real counterpart of preconditioners.rs:934-1010): kind 0 none, 1 Jacobi
    (preconditioners.rs:210-229), 3 given packed LU factors with perm (P A = L U,
    perm[i] the original row at position i).  (x, iterations, residual_norm,
    converged)."""
    a = _c(a, np.float64)
    b = _c(b, np.float64)
    n = a.shape[0]
    x = np.empty(n)
    it = np.zeros(1, dtype=np.int32)
    conv = np.zeros(1, dtype=np.int32)
    res = np.zeros(1)
    x0p = None if x0 is None else _c(x0, np.float64)
    lup = None if lu is None else _c(lu, np.float64)
    pp = None if perm is None else np.ascontiguousarray(perm, dtype=np.int64)
    rc = _load().oracle_pgmres_d(n, _ptr(a), n, _ptr(b), None if x0p is None else _ptr(x0p), int(max_iter),
                                 float(tol), int(restart), int(kind), None if lup is None else _ptr(lup),
                                 None if pp is None else _ptr(pp), _ptr(x), _ptr(it), _ptr(res), _ptr(conv))
    if rc > 0:
        raise OracleError("singular", rc)
    if rc < 0:
        raise OracleError("argument" if rc == -2 else "memory", rc)
    return x, int(it[0]), float(res[0]), bool(conv[0])
