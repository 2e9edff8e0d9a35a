"""Right-preconditioned restarted GMRES (scirs_gpu_dgmres_precond), the real
counterpart of scirs2-linalg/src/preconditioners.rs:934-1010, against the
oracle's restatement of the gmres iteration (solvers/iterative.rs:317-450) with
right preconditioning (oracle.pgmres).  With the same preconditioner both sides
take the same steps, so the iteration count and the convergence flag must
match and the iterates agree to rounding (bar: 1e-9 kappa relative).  The
library-built LU preconditioner has no oracle twin (its pivots come from the
tournament panel): it is checked against LAPACK's solution and against the
unpreconditioned iteration count."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from scirs_amd import linalg
    return linalg


@pytest.fixture(scope="module")
def N():
    from scirs_amd import _native
    return _native


def row_scaled(n, seed, spread=3):
    rng = np.random.default_rng(seed)
    d = np.logspace(-spread, spread, n)
    a = d[:, None] * (np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n))
    return a, rng.standard_normal(n)


# (n, restart, decades of row scaling, tol): cases whose last two cycles end
# far on either side of tol (1.3e-7 -> 1.3e-11, ..., 1.6e-7 -> 3.3e-12 in the
# oracle), so rounding differences cannot move the stopping cycle
@pytest.mark.parametrize("n,restart,spread,tol", [(30, 10, 3, 1e-10), (120, 10, 2, 1e-9), (300, 20, 2, 1e-8)])
def test_jacobi_matches_oracle(L, N, n, restart, spread, tol):
    a, b = row_scaled(n, n, spread)
    print(f"pgmres jacobi n={n} restart={restart}", flush=True)
    x, it, res, conv = L.gmres_precond(a, b, N.PRECOND_DIAGONAL, max_iter=2000, tol=tol, restart=restart)
    xo, ito, reso, convo = oracle.pgmres(a, b, 1, max_iter=2000, tol=tol, restart=restart)
    assert (it, conv) == (ito, convo)
    kappa = np.linalg.cond(a)
    assert np.abs(x - xo).max() <= 1e-9 * kappa * max(1.0, np.abs(xo).max())
    assert conv and res < tol


@pytest.mark.parametrize("n", [64, 257])
def test_given_lu_matches_oracle(L, N, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    lu32, piv32, _ = oracle.lu_factor(a.astype(np.float32))
    lu = lu32.astype(np.float64)
    print(f"pgmres given LU n={n}", flush=True)
    x, it, res, conv = L.gmres_precond(a, b, N.PRECOND_LU_GIVEN, lu=lu, perm=piv32, max_iter=300, tol=1e-11,
                                       restart=10)
    xo, ito, reso, convo = oracle.pgmres(a, b, 3, lu=lu, perm=piv32, max_iter=300, tol=1e-11, restart=10)
    assert (it, conv) == (ito, convo) and conv
    assert np.abs(x - xo).max() <= 1e-9 * np.linalg.cond(a) * max(1.0, np.abs(xo).max())


def test_given_lu_after_a_breakdown(L, N):
    """A singular solve leaves its breakdown column in the library's flag; the
    given-LU preconditioner's triangular solves must not read it as their own."""
    s = np.ones((6, 6))
    with pytest.raises(L.SingularError):
        L.gesv(s, np.ones(6))
    n = 64
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    lu32, piv32, _ = oracle.lu_factor(a.astype(np.float32))
    lu = lu32.astype(np.float64)
    print("pgmres given LU after a singular gesv", flush=True)
    x, it, res, conv = L.gmres_precond(a, b, N.PRECOND_LU_GIVEN, lu=lu, perm=piv32, max_iter=300, tol=1e-11,
                                       restart=10)
    _, ito, _, convo = oracle.pgmres(a, b, 3, lu=lu, perm=piv32, max_iter=300, tol=1e-11, restart=10)
    assert (it, conv) == (ito, convo) and conv


@pytest.mark.parametrize("n", [500, 2000])
def test_library_lu_preconditioner(L, N, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 0.5 * np.eye(n)
    b = rng.standard_normal(n)
    print(f"pgmres library LU n={n}", flush=True)
    x, it, res, conv = L.gmres_precond(a, b, N.PRECOND_LU, max_iter=600, tol=1e-10, restart=20)
    assert conv and res < 1e-10
    xs = np.linalg.solve(a, b)
    assert np.abs(x - xs).max() <= 1e-9 * np.linalg.cond(a) * max(1.0, np.abs(xs).max())
    _, it0, _, c0 = L.gmres(a, b, max_iter=600, tol=1e-10, restart=20)
    assert it <= 40 and (not c0 or it < it0), (it, it0, c0)


def test_none_is_plain_gmres(L, N):
    rng = np.random.default_rng(8)
    n = 100
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 2.0 * np.eye(n)
    b = rng.standard_normal(n)
    x, it, res, conv = L.gmres_precond(a, b, N.PRECOND_NONE)
    xg, itg, resg, convg = L.gmres(a, b)
    assert (it, conv) == (itg, convg) and np.array_equal(x, xg)


def test_invalid_arguments(L, N):
    a = np.eye(4)
    with pytest.raises(L.GpuSolveError):
        L.gmres_precond(a, np.ones(4), N.PRECOND_LU_GIVEN)          # factors missing
    with pytest.raises(L.GpuSolveError):
        L.gmres_precond(a, np.ones(4), 7)                            # unknown kind
    with pytest.raises(L.GpuSolveError):
        L.gmres_precond(a, np.ones(4), N.PRECOND_LU_GIVEN, lu=a, perm=np.array([0, 1, 2, 9]))
