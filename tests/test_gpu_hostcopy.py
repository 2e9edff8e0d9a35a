"""Host-memory entry points on strided host arrays (leading dimension larger
than the row), large enough that the pinned staging between host and device
runs in several chunks and reuses both of its buffers.

Bar: the factors and solutions equal, bit for bit, those of the same call on
contiguous copies of the same data (the device computation is the same; only
the transfer differs), every element outside the strided rows is left as it
was, and the solution meets the FP64 backward-error bar 8 n 2^-53.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U64 = 2.0 ** -53


@pytest.fixture(scope="module")
def N():
    from scirs_amd import _native
    _native.load()
    return _native


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def strided(rng, rows, cols, pad):
    """rows x cols view (leading dimension cols + pad) of a NaN-padded array."""
    full = np.full((rows, cols + pad), np.nan)
    full[:, :cols] = rng.standard_normal((rows, cols))
    return full


def bwd(a, x, b):
    r = b - a @ x
    na = np.abs(a).sum(axis=1).max()
    return max(np.abs(r[:, j]).max() / (na * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(x.shape[1]))


@pytest.mark.parametrize("n,pad", [(5, 3), (2100, 77), (4100, 13)])
def test_dgetrf_strided_matches_contiguous(N, n, pad):
    lib = N.load()
    rng = np.random.default_rng(n + pad)
    full = strided(rng, n, n, pad)
    dense = np.ascontiguousarray(full[:, :n])
    perm_s = np.empty(n, np.int64)
    perm_c = np.empty(n, np.int64)
    info = N.SolveInfo()
    print(f"dgetrf strided n={n} lda={n + pad}", flush=True)
    assert lib.scirs_gpu_dgetrf(n, n, _p(full), n + pad, _p(perm_s), ctypes.byref(info)) == 0
    assert lib.scirs_gpu_dgetrf(n, n, _p(dense), n, _p(perm_c), ctypes.byref(info)) == 0
    assert np.array_equal(perm_s, perm_c)
    assert np.array_equal(full[:, :n], dense)
    assert np.isnan(full[:, n:]).all()


@pytest.mark.parametrize("n,nrhs", [(7, 2), (2100, 3), (3000, 1)])
def test_dgesv_strided_matches_contiguous(N, n, nrhs):
    lib = N.load()
    rng = np.random.default_rng(10 * n + nrhs)
    af = strided(rng, n, n, 31)
    bf = strided(rng, n, nrhs, 5)
    xf = np.full((n, nrhs + 4), np.nan)
    a = np.ascontiguousarray(af[:, :n])
    b = np.ascontiguousarray(bf[:, :nrhs])
    x = np.empty((n, nrhs))
    info = N.SolveInfo()
    print(f"dgesv strided n={n} nrhs={nrhs}", flush=True)
    assert lib.scirs_gpu_dgesv(n, nrhs, _p(af), n + 31, _p(bf), nrhs + 5, _p(xf), nrhs + 4,
                               ctypes.byref(info)) == 0
    assert lib.scirs_gpu_dgesv(n, nrhs, _p(a), n, _p(b), nrhs, _p(x), nrhs, ctypes.byref(info)) == 0
    assert np.array_equal(xf[:, :nrhs], x)
    assert np.isnan(xf[:, nrhs:]).all()
    assert bwd(a, x, b) <= 8 * n * U64


def test_sgesv_strided_matches_contiguous(N):
    lib = N.load()
    n, nrhs = 2300, 2
    rng = np.random.default_rng(77)
    af = np.full((n, n + 9), np.nan, np.float32)
    af[:, :n] = rng.standard_normal((n, n))
    bf = np.full((n, nrhs + 1), np.nan, np.float32)
    bf[:, :nrhs] = rng.standard_normal((n, nrhs))
    xf = np.full((n, nrhs + 3), np.nan, np.float32)
    a = np.ascontiguousarray(af[:, :n])
    b = np.ascontiguousarray(bf[:, :nrhs])
    x = np.empty((n, nrhs), np.float32)
    info = N.SolveInfo()
    print(f"sgesv strided n={n}", flush=True)
    assert lib.scirs_gpu_sgesv(n, nrhs, _p(af), n + 9, _p(bf), nrhs + 1, _p(xf), nrhs + 3, ctypes.byref(info)) == 0
    assert lib.scirs_gpu_sgesv(n, nrhs, _p(a), n, _p(b), nrhs, _p(x), nrhs, ctypes.byref(info)) == 0
    assert np.array_equal(xf[:, :nrhs], x)
    assert np.isnan(xf[:, nrhs:]).all()


def test_batch_matmul_flat_copy_with_tail(N):
    """A batch whose bytes are not a multiple of the 64 KiB rows of the flat
    copy: the tail row travels too (the result is exact on small integers)."""
    from scirs_amd import linalg
    rng = np.random.default_rng(3)
    a = rng.integers(-4, 5, (3, 101, 77)).astype(np.float64)
    b = rng.integers(-4, 5, (77, 13)).astype(np.float64)
    print("batch_matmul 3x101x77 @ 77x13", flush=True)
    out = linalg.batch_matmul(a, b)
    assert np.array_equal(out, np.einsum("pik,kj->pij", a, b))
