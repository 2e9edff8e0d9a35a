"""The measured configurations at their real sizes, through size-independent
properties (the oracle is too slow at these sizes; it pins the same code paths
at n <= 1536 in test_gpu_lu / test_gpu_mixed / test_gpu_batched).

* n = 8192, 16 right-hand sides, plain FP64: LU solve of a N(0,1) system and
  Cholesky solve of G G^T / n + I (the reference's `solve_multiple` and the
  cholesky + two substitutions composition).  Bar: normwise backward error per
  column <= 8 n 2^-53; the solution agrees with LAPACK's (numpy) within
  8 n 2^-53 kappa_inf.
* n = 16384, kappa_2 = 1e4 (A = U diag(logspace) V^T), mixed-precision IR:
  backward error <= 1e-14 and <= n 2^-53 on the mixed path; forward error
  against x_true <= 100 kappa_2 n 2^-53.
* 20,000 uneven systems (n ~ U[8, 256], half SPD) in one batched launch: every
  system solved (status 0) with backward error <= 8 n_i 2^-53, and a sample
  of 40 systems equal to the oracle's solution within 8 n 2^-53 kappa_inf.
* n = 32768, kappa_2 = 1e8, 8 right-hand sides, FP16 update: FP64 backward
  error <= n 2^-53 (GMRES-IR, or the FP64 fallback when GMRES cannot reach it).
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


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


def eta_cols(a, x, b):
    r = b - a @ x
    na = np.abs(a).sum(axis=1).max()
    return np.abs(r).max(axis=0) / (na * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))


def conditioned_gpu(torch, n, kappa, nrhs, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    u, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    v, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    sigma = torch.logspace(0.0, float(np.log10(kappa)), n, dtype=torch.float64, device="cuda")
    a = ((u * sigma) @ v.T).contiguous()
    del u, v
    x_true = torch.randn(n, nrhs, dtype=torch.float64, device="cuda", generator=g)
    b = (a @ x_true).contiguous()
    torch.cuda.synchronize()
    return a, b, x_true


def test_config_lu_and_cholesky_n8192_16rhs(L):
    n, nrhs = 8192, 16
    rng = np.random.default_rng(8192)
    a = rng.standard_normal((n, n))
    b = rng.standard_normal((n, nrhs))
    print("config: LU solve n=8192 nrhs=16", flush=True)
    x, _ = L.gesv(a, b)
    assert eta_cols(a, x, b).max() <= 8 * n * U64
    xl = np.linalg.solve(a, b)
    kappa = np.linalg.norm(a, np.inf) * np.linalg.norm(np.linalg.inv(a), np.inf)
    assert np.abs(x - xl).max() / np.abs(xl).max() <= 8 * n * U64 * kappa
    del xl
    g = rng.standard_normal((n, n))
    s = g @ g.T / n + np.eye(n)
    del g
    print("config: Cholesky solve n=8192 nrhs=16", flush=True)
    xc, _ = L.posv(s, b)
    assert eta_cols(s, xc, b).max() <= 8 * n * U64
    xl = np.linalg.solve(s, b)
    kappa = np.linalg.norm(s, np.inf) * np.linalg.norm(np.linalg.inv(s), np.inf)
    assert np.abs(xc - xl).max() / np.abs(xl).max() <= 8 * n * U64 * kappa


def test_config_mixed_n16384_kappa1e4(L, torch):
    n, kappa = 16384, 1e4
    print("config: mixed IR n=16384 kappa=1e4", flush=True)
    a, b, x_true = conditioned_gpu(torch, n, kappa, 1, 2024)
    x = torch.empty_like(b)
    info = L.gesv_ir_dev(a, b, x, tol=1e-14)
    eta = L.backward_error_dev(a, x, b)
    assert eta <= min(1e-14, n * U64), (eta, info.path, info.ir_iters, info.gmres_iters)
    assert info.path in (2, 3)
    ferr = float((x - x_true).abs().max() / x_true.abs().max())
    assert ferr <= 100 * kappa * n * U64


def test_config_batched_20000(L, torch):
    import bench
    print("config: batched 20000 systems n~U[8,256]", flush=True)
    n, nrhs, spd, offA, offB, dA, dB = bench.make_batch(20000, 2024, torch)
    dX = torch.zeros_like(dB)
    status, bwd, info = L.gesv_batched_dev(n, nrhs, offA, offB, dA, dB, dX, spd)
    assert (status == 0).all(), np.nonzero(status)[0][:10]
    assert (bwd <= 8 * n * U64).all()
    assert info.failed == 0
    A = dA.cpu().numpy()
    B = dB.cpu().numpy()
    X = dX.cpu().numpy()
    rng = np.random.default_rng(5)
    for i in rng.choice(len(n), 40, replace=False):
        m, r = int(n[i]), int(nrhs[i])
        a = A[offA[i]:offA[i] + m * m].reshape(m, m)
        b = B[offB[i]:offB[i] + m * r].reshape(m, r)
        x = X[offB[i]:offB[i] + m * r].reshape(m, r)
        xo = oracle.cholesky_solve(a, b) if spd[i] else oracle.solve_multiple(a, b)
        kappa = np.linalg.cond(a, np.inf)
        assert np.abs(x - xo).max() / np.abs(xo).max() <= 8 * m * U64 * kappa, (i, m, int(spd[i]))


def test_config_hard_fp16_n32768_kappa1e8(L, torch):
    """The hard case: FP16x3 factors (FP16 matrix cores, scaled hi/lo split),
    GMRES-IR with the preconditioner applied in FP64, 8 right-hand sides side
    by side.  With sigma log-spaced over [1, 1e8], every factorisation below
    FP64 leaves a fixed fraction of the singular directions unresolved (those
    with sigma_i below about u_f * growth * sigma_max), so GMRES needs a number
    of iterations proportional to n: measured, 120 per right-hand side at
    n = 1024, 990 at n = 2048 (restart 600), more than 2000 at n = 4096 even
    with FP32-accurate factors.  At n = 32768 the solve therefore ends on the
    FP64 fallback after GMRES stalls; the result must still meet the FP64 bar,
    and the test records the path (expected failure while it is 4)."""
    n, nrhs, kappa = 32768, 8, 1e8
    print("config: FP16x3 GMRES-IR n=32768 kappa=1e8 nrhs=8", flush=True)
    a, b, x_true = conditioned_gpu(torch, n, kappa, nrhs, 32768)
    x = torch.empty_like(b)
    info = L.gesv_ir_dev(a, b, x, mode=6)
    eta = L.backward_error_dev(a, x, b)
    print("path", info.path, "ir", info.ir_iters, "gmres", info.gmres_iters, "eta", eta, flush=True)
    assert eta <= n * U64
    assert info.path in (3, 4)
    if info.path == 4:
        pytest.xfail(f"GMRES-IR stalled after {info.gmres_iters} iterations; FP64 LU answered "
                     "(iterations grow with n for this spectrum)")
