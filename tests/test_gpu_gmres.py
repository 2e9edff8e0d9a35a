"""Parity of the HIP GMRES with the oracle (solvers/iterative.rs:317-450).

Bar: the same iteration count and convergence flag as the oracle, x within
1e-9 (relative to max|x|) of the oracle's, the reported residual norm equal to
||b - A x||_2 recomputed on the host within 1e-12 + 1e-9 relative (the Krylov
basis is formed with sums in a different order, so iterates agree to rounding,
not bit for bit).
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from scirs_amd import linalg
    return linalg


def _check(L, a, b, **kw):
    x, it, res, conv = L.gmres(a, b, **kw)
    xo, ito, reso, convo = oracle.gmres(a, b, **kw)
    assert (it, conv) == (ito, convo)
    assert np.abs(x - xo).max() <= 1e-9 * max(1.0, np.abs(xo).max())
    rn = np.linalg.norm(b - a @ x)
    assert abs(res - rn) <= 1e-12 + 1e-9 * rn
    return x, it, res, conv


def test_reference_example(L):
    x, it, res, conv = _check(L, np.array([[1.0, 2.0], [3.0, 4.0]]), np.array([5.0, 11.0]))
    assert conv and np.allclose(x, [1.0, 2.0], atol=1e-10)


@pytest.mark.parametrize("n", [5, 33, 200, 1000])
def test_matches_oracle(L, n):
    print("gmres n", n, flush=True)
    rng = np.random.default_rng(n + 1)
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 3.0 * np.eye(n)
    b = rng.standard_normal(n)
    _check(L, a, b)


def test_restarts_max_iter_and_x0(L):
    rng = np.random.default_rng(4)
    n = 300
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 1.5 * np.eye(n)
    b = rng.standard_normal(n)
    _check(L, a, b, max_iter=5)                    # one cycle, 5 steps, not converged
    _check(L, a, b, restart=10, max_iter=200)      # many restarts
    _check(L, a, b, x0=np.linalg.solve(a, b))      # converged before any step
    _check(L, a, b, max_iter=0)


def test_breakdown_ends_the_cycle(L):
    # A = I and b = 2 e_3: v_0 = e_3 and w = v_0 exactly, so h_10 = 0 < eps at
    # the first step (counted as no iteration)
    n = 50
    b = np.zeros(n)
    b[3] = 2.0
    x, it, res, conv = _check(L, np.eye(n), b)
    assert conv and it == 0


def test_dev_entry_matches_host_entry(L):
    """The device-resident entry point runs the same kernels: same iterations,
    same bits of x, with and without x0 (lda larger than n covered)."""
    import torch
    rng = np.random.default_rng(12)
    n = 500
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 2.0 * np.eye(n)
    b = rng.standard_normal(n)
    x_h, it_h, res_h, conv_h = L.gmres(a, b, restart=20)
    big = np.zeros((n, n + 36))
    big[:, :n] = a
    ta = torch.from_numpy(big).cuda()[:, :n]
    tb = torch.from_numpy(b).cuda()
    tx = torch.full((n,), 7.0, dtype=torch.float64, device="cuda")
    it, res, conv = L.gmres_dev(ta, tb, tx, restart=20)
    assert (it, conv) == (it_h, conv_h)
    assert np.array_equal(tx.cpu().numpy(), x_h) and res == res_h
    x0 = np.linalg.solve(a, b)
    tx = torch.from_numpy(x0).cuda()
    it, res, conv = L.gmres_dev(ta, tb, tx, use_x0=True)
    assert it == 0 and conv
