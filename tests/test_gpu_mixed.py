"""Mixed-precision solve (scirs_gpu_dgesv_ir): low-precision LU + FP64
refinement, GMRES-IR fallback.  The reference's counterpart is
scirs2-linalg/src/mixed_precision/adaptive.rs:282-378 (refinement) and
solvers/iterative.rs:317-450 (GMRES); the result must be the FP64 solution.

Bars (floating point, written per test):
* normwise backward error eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)
  <= n * 2^-53, the solver's own stopping test;
* forward error vs. the oracle's FP64 solution and vs. x_true
  <= 100 * kappa_2 * n * 2^-53 (kappa_2 is exact by construction);
* the solve stays on the mixed path (refinement or GMRES-IR), no FP64 fallback,
  for the modes whose unit roundoff times kappa is below ~5 (FP32, BF16x3, FP16
  with scaling); a pure BF16 update at kappa = 1e4 (kappa * 2^-8 ~ 39) may need
  the FP64 fallback, and then only the accuracy bars apply.
The four trailing-update arithmetics are all exercised, on leading dimensions
that are and are not multiples of 4 (vector and scalar loads of the update) and
on sizes that leave ragged 128-row tiles.
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

U64 = 2.0 ** -53
MODES = {"bf16": 1, "fp16": 2, "bf16x3": 3, "fp32": 4, "bf16x6": 5, "fp16x3": 6}


@pytest.fixture(scope="module")
def L():
    from scirs_amd import linalg
    return linalg


def conditioned(n, kappa, seed, nrhs=1):
    """A = U diag(sigma) V^T with sigma log-spaced over [1, kappa]; b = A x_true."""
    rng = np.random.default_rng(seed)
    u, _ = np.linalg.qr(rng.standard_normal((n, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sigma = np.logspace(0.0, np.log10(kappa), n)
    a = (u * sigma) @ v.T
    x_true = rng.standard_normal((n, nrhs))
    return a, a @ x_true, x_true


def eta(a, x, b):
    x = x.reshape(a.shape[0], -1)
    b = b.reshape(a.shape[0], -1)
    r = b - a @ x
    na = np.abs(a).sum(axis=1).max()
    return max(np.abs(r[:, j]).max() / (na * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(x.shape[1]))


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("n", [333, 1000, 2048])
def test_mixed_modes_reach_fp64(L, mode, n):
    kappa = 1e4
    a, b, x_true = conditioned(n, kappa, seed=n + MODES[mode])
    print(f"mixed mode={mode} n={n}", flush=True)
    x, info = L.gesv_ir(a, b, mode=MODES[mode])
    e = eta(a, x, b)
    detail = (mode, n, e, info.path, info.ir_iters, info.gmres_iters)
    assert e <= n * U64, detail
    if mode != "bf16":
        assert info.path in (2, 3), f"left the mixed path: {detail}"
    print("path", info.path, "ir", info.ir_iters, "gmres", info.gmres_iters, flush=True)
    assert info.backward_error <= n * U64
    bound = 100 * kappa * n * U64
    assert np.abs(x - x_true).max() / np.abs(x_true).max() <= bound
    xo = oracle.solve_multiple(a, b)
    assert np.abs(x - xo).max() / np.abs(xo).max() <= bound


@pytest.mark.parametrize("n", [130, 700, 1300])
def test_refinement_links_with_odd_block_count(L, n):
    """The one-rhs FP32 refinement solves hand over two 64-row block rows per
    link; an odd number of blocks (3, 11, 21) leaves the last link one.  A wrong
    link would stall the refinement: the solve must stay on plain refinement
    (path 2) in as few steps as on even counts, and meet the FP64 bars."""
    kappa = 1e4
    a, b, x_true = conditioned(n, kappa, seed=7 * n)
    print(f"mixed bf16x6 n={n} (odd block count)", flush=True)
    x, info = L.gesv_ir(a, b, mode=MODES["bf16x6"])
    detail = (n, info.path, info.ir_iters, info.backward_error)
    assert info.path == 2 and info.ir_iters <= 6, detail
    assert eta(a, x, b) <= n * U64, detail
    assert np.abs(x - x_true).max() / np.abs(x_true).max() <= 100 * kappa * n * U64


def test_mixed_multiple_rhs(L):
    n, nrhs, kappa = 777, 5, 1e3
    a, b, x_true = conditioned(n, kappa, seed=5, nrhs=nrhs)
    x, info = L.gesv_ir(a, b, mode=MODES["bf16x3"])
    assert x.shape == (n, nrhs)
    assert eta(a, x, b) <= n * U64
    assert np.abs(x - x_true).max() / np.abs(x_true).max() <= 100 * kappa * n * U64


def test_hard_case_goes_through_gmres(L):
    """kappa = 1e8: plain refinement cannot converge, GMRES-IR (FP64 Arnoldi,
    preconditioner applied in FP64 on the factors upcast once) must, without
    the FP64 fallback.  The factors come from the FP16x3 update (FP16 matrix
    cores, power-of-two scaled hi/lo split, three products): at n = 1024 they
    leave few enough of the spectrum's directions unresolved for GMRES(50)
    within its budget (about 120 iterations per right-hand side on this
    seed)."""
    n, nrhs, kappa = 1024, 2, 1e8
    a, b, x_true = conditioned(n, kappa, seed=8, nrhs=nrhs)
    x, info = L.gesv_ir(a, b, mode=MODES["fp16x3"])
    e = eta(a, x, b)
    detail = (e, info.path, info.ir_iters, info.gmres_iters)
    assert e <= n * U64, detail
    assert info.path == 3 and info.gmres_iters > 0, detail
    assert np.abs(x - x_true).max() / np.abs(x_true).max() <= 100 * kappa * n * U64


def test_gmres_ir_beyond_its_reach_falls_back(L):
    """One FP16 product at kappa = 1e8 (u_16 kappa ~ 5e4): GMRES-IR stalls
    (two restart cycles without a 10 % residual reduction), and FP64 LU
    answers (path 4) to FP64 accuracy."""
    n, nrhs, kappa = 1024, 2, 1e8
    a, b, x_true = conditioned(n, kappa, seed=8, nrhs=nrhs)
    x, info = L.gesv_ir(a, b, mode=MODES["fp16"])
    assert info.path == 4 and info.gmres_iters > 0, (info.path, info.gmres_iters)
    assert eta(a, x, b) <= n * U64


def test_mixed_singular_reports_column(L):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((200, 200))
    a[:, 57] = 0.0
    with pytest.raises(L.SingularError) as e:
        L.gesv_ir(a, np.ones(200), mode=MODES["bf16x3"])
    # the FP64 breakdown column of the reference (lapack.rs:106-120)
    with pytest.raises(oracle.OracleError) as eo:
        oracle.lu_factor(a)
    assert e.value.rc == eo.value.code


@pytest.mark.parametrize("n,nrhs", [(2, 1), (3, 1), (333, 1), (1000, 1), (1000, 3), (257, 5)])
def test_backward_error_matches_host(L, n, nrhs):
    """The device normwise backward error (FP64 residual, norms) against numpy,
    for an x that is not a solution: one right-hand side (16-byte vector loads,
    odd and even n) and several (the general residual kernel)."""
    import torch
    rng = np.random.default_rng(n + nrhs)
    a = rng.standard_normal((n, n))
    x = rng.standard_normal((n, nrhs))
    b = rng.standard_normal((n, nrhs))
    ta, tx, tb = (torch.from_numpy(v).cuda() for v in (a, x, b))
    got = L.backward_error_dev(ta, tx, tb)
    want = eta(a, x, b)
    assert abs(got - want) <= 1e-12 * want, (got, want)
