"""Tournament-pivoted FP32 LU (scirs_gpu_sgetrf_tp_dev): the factorisation the
mixed-precision solve refines.  Its pivots come from a reduction tree over
blocks of rows (CALU), not from the column arg-max of
scirs2-linalg/src/lapack.rs:95-104, so it is checked by what any LU must
satisfy, computed on the host in FP64:

* perm is a permutation of 0..n-1;
* P A = L U: max |P A - L U| <= 8 n 2^-24 max(|L| |U|) (FP32 arithmetic; the
  BF16x6 update is FP32-accurate, so the same bar);
* L is unit lower triangular with entries bounded (tournament pivoting keeps
  |l_ij| small in practice; the bar is 32);
* an exactly zero column k stops the factorisation with status k + 1, the
  absolute test |u_kk| < eps of lapack.rs:106-120.

Sizes cover a single workgroup (n <= 1024 rows per sub-panel), a two-level
tree, ragged sub-panels (n = 333 leaves a 13-column sub-panel) and row strides
that are not multiples of 64.  At n = 33000 the first panels have 33 blocks of
1024 rows (a three-level tree whose last block is its parent's sole child);
there P A = L U is checked on random vectors (O(n^2) work, in FP64 with torch
on the GPU).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
MODE_FP32, MODE_BF16X6 = 4, 5


@pytest.fixture(scope="module")
def L():
    from scirs_amd import linalg
    return linalg


def factor(L, a, pad=4):
    import torch
    n = a.shape[0]
    ld = (n + 3) // 4 * 4 + pad
    buf = torch.zeros((n, ld), dtype=torch.float32, device="cuda")
    buf[:, :n] = torch.from_numpy(a)
    A = buf[:, :n]
    perm = torch.empty(n, dtype=torch.int64, device="cuda")
    return A, perm


@pytest.mark.parametrize("mode", [MODE_FP32, MODE_BF16X6])
@pytest.mark.parametrize("n", [8, 100, 333, 512, 1000, 1537, 4200, 4700])
def test_tournament_lu_factors(L, n, mode):
    if mode == MODE_BF16X6 and n not in (333, 1000, 4700):
        pytest.skip("BF16x6 update checked on three sizes")
    rng = np.random.default_rng(n + 17 * mode)
    a = rng.standard_normal((n, n)).astype(np.float32)
    print(f"tslu n={n} mode={mode}", flush=True)
    A, perm = factor(L, a, pad=4 + 4 * (n % 3))
    rc = L.sgetrf_tp_dev(A, perm, mode)
    assert rc == 0
    lu = A.cpu().numpy().astype(np.float64)
    p = perm.cpu().numpy()
    assert np.array_equal(np.sort(p), np.arange(n)), "perm is not a permutation"
    low = np.tril(lu, -1) + np.eye(n)
    up = np.triu(lu)
    growth_l = np.abs(low).max()
    assert growth_l <= 32, growth_l
    r = a.astype(np.float64)[p] - low @ up
    scale = (np.abs(low) @ np.abs(up)).max()
    err = np.abs(r).max() / scale
    assert err <= 8 * n * U32, (err, growth_l)


def test_tournament_lu_zero_column(L):
    n = 700
    rng = np.random.default_rng(3)
    a = rng.standard_normal((n, n)).astype(np.float32)
    a[:, 57] = 0.0
    A, perm = factor(L, a)
    assert L.sgetrf_tp_dev(A, perm, MODE_FP32) == 58


def test_tournament_lu_rejects_unaligned_stride(L):
    import torch
    A = torch.zeros((10, 11), dtype=torch.float32, device="cuda")[:, :10]
    perm = torch.empty(10, dtype=torch.int64, device="cuda")
    with pytest.raises(L.GpuSolveError):
        L.sgetrf_tp_dev(A, perm, MODE_FP32)


def test_tournament_lu_three_levels(L):
    import torch
    n = 33000
    g = torch.Generator(device="cuda").manual_seed(12)
    a = torch.randn(n, n, dtype=torch.float32, device="cuda", generator=g)
    A = a.clone()
    perm = torch.empty(n, dtype=torch.int64, device="cuda")
    print(f"tslu n={n} (three-level tree)", flush=True)
    assert L.sgetrf_tp_dev(A, perm, MODE_FP32) == 0
    p = perm.cpu().numpy()
    assert np.array_equal(np.sort(p), np.arange(n)), "perm is not a permutation"
    x = torch.randn(n, 4, dtype=torch.float64, device="cuda", generator=g)
    low = torch.tril(A, -1).double()
    low.diagonal().fill_(1.0)
    up = torch.triu(A).double()
    assert low.abs().max().item() <= 32
    lhs = a.double()[perm] @ x
    rhs = low @ (up @ x)
    scale = (low.abs() @ (up.abs() @ x.abs())).max().item()
    err = (lhs - rhs).abs().max().item() / scale
    assert err <= 8 * n * U32, err
