"""The allocation-free entry point scirs_gpu_backward_error_dev_noalloc: its
scratch is a fixed array in the code object.  Its result must equal the
allocating entry point's (same kernels, same order: bit-identical) on two calls
in a row, the second smaller than the first (the array keeps the first call's
contents, so nothing of it may be read before it is written), and it must
refuse sizes beyond the array and a library that was not initialised."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from scirs_amd import linalg
    return linalg


def _case(torch, n, nrhs, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    x = rng.standard_normal((n, nrhs))
    b = rng.standard_normal((n, nrhs))   # not A x: no cancellation in the torch reference
    return tuple(torch.from_numpy(v).cuda() for v in (a, x, b))


def test_noalloc_matches_allocating_entry_point(L):
    import torch
    from scirs_amd import _native as N
    assert N.load().scirs_gpu_init(0) == 0
    for n, nrhs, seed in ((3000, 5, 1), (500, 2, 2)):
        print(f"noalloc backward error n={n} nrhs={nrhs}", flush=True)
        A, X, B = _case(torch, n, nrhs, seed)
        want = L.backward_error_dev(A, X, B)
        got = L.backward_error_dev_noalloc(A, X, B)
        assert got == want, (n, nrhs, got, want)
        r = (B - A @ X).abs().max(dim=0).values
        den = A.abs().sum(dim=1).max() * X.abs().max(dim=0).values + B.abs().max(dim=0).values
        ref = float((r / den).max())
        assert abs(got - ref) <= 1e-12 * ref


def test_noalloc_rejects_sizes_beyond_its_scratch(L):
    import torch
    from scirs_amd import _native as N
    lib = N.load()
    assert lib.scirs_gpu_init(0) == 0
    n, nrhs = 1024, 1024   # n * nrhs + n + 200 > 2^20 doubles
    A = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    X = torch.zeros((n, nrhs), dtype=torch.float64, device="cuda")
    out = ctypes.c_double(0.0)
    rc = lib.scirs_gpu_backward_error_dev_noalloc(n, nrhs, ctypes.c_void_p(A.data_ptr()), n,
                                                  ctypes.c_void_p(X.data_ptr()), nrhs,
                                                  ctypes.c_void_p(X.data_ptr()), nrhs, None, ctypes.byref(out))
    assert rc == N.SCIRS_ERR_ARG
    assert "exceed its fixed scratch" in N.last_error()
