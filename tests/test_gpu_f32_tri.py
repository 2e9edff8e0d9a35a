"""f32 triangular solve (scirs_gpu_strsv) and f32 Cholesky solve
(scirs_gpu_sposv) against the oracle's f32 instantiations of
scirs2-linalg/src/solve.rs:139-202 and lapack.rs:632-679.

Both sides compute in f32 with different summation orders, so the bar is a
forward bound, |x - x_oracle| <= 8 n 2^-24 kappa_inf(A) max|x_oracle|, plus
the f32 normwise backward error eta <= 8 n 2^-24.  Two checks show the f32
semantics themselves, which an f64 computation rounded to f32 would fail:

* a diagonal of 1e-8 is singular for f32 (below f32::EPSILON, solve.rs:166)
  and not for f64;
* on a long substitution the f32 result carries rounding error of several f32
  ulps against the exact (f64) solution, like the oracle's f32 loop, where an
  f64 solve rounded once to f32 stays within half an ulp.
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


@pytest.fixture(scope="module")
def L():
    from scirs_amd import linalg
    return linalg


def tri(rng, n, lower):
    a = rng.standard_normal((n, n)) / np.sqrt(n)
    a = np.tril(a) if lower else np.triu(a)
    np.fill_diagonal(a, 1.0 + rng.random(n))
    return a.astype(np.float32)


def bwd32(a, x, b):
    a = np.asarray(a, np.float64)
    x = np.asarray(x, np.float64).reshape(a.shape[0], -1)
    b = np.asarray(b, np.float64).reshape(a.shape[0], -1)
    r = b - a @ x
    na = np.abs(a).sum(axis=1).max()
    return max(np.abs(r[:, j]).max() / (na * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(x.shape[1]))


@pytest.mark.parametrize("n", [1, 64, 333, 1000])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_strsv_matches_oracle_f32(L, n, nrhs):
    rng = np.random.default_rng(31 * n + nrhs)
    for lower in (True, False):
        for unit in (True, False):
            a = tri(rng, n, lower)
            b = rng.standard_normal((n, nrhs)).astype(np.float32)
            print(f"strsv n={n} nrhs={nrhs} lower={lower} unit={unit}", flush=True)
            x, _ = L.trsv(a, b, lower, unit, dtype=np.float32)
            assert x.dtype == np.float32
            a_eff = a.astype(np.float64)
            if unit:
                np.fill_diagonal(a_eff, 1.0)
            kappa = np.linalg.cond(a_eff, np.inf)
            for j in range(nrhs):
                xo = oracle.solve_triangular(a, np.ascontiguousarray(b[:, j]), lower, unit)
                assert xo.dtype == np.float32
                err = np.abs(x[:, j].astype(np.float64) - xo).max() / np.abs(xo).max()
                assert err <= 8 * n * U32 * kappa, (lower, unit, j, err)
            assert bwd32(a_eff, x, b) <= 8 * n * U32


def test_strsv_uses_f32_epsilon(L):
    a = np.array([[1.0, 0.0, 0.0], [2.0, 1e-8, 0.0], [1.0, 1.0, 1.0]])
    with pytest.raises(L.SingularError) as eg:
        L.trsv(a, np.ones(3), True, False, dtype=np.float32)
    assert eg.value.rc == 2
    with pytest.raises(oracle.OracleError) as eo:
        oracle.solve_triangular(a.astype(np.float32), np.ones(3, np.float32), True, False)
    assert eo.value.code == 2
    x, _ = L.trsv(a, np.ones(3), True, False)   # f64: 1e-8 is a valid pivot
    assert np.isfinite(x).all()


def test_strsv_rounds_like_f32_substitution(L):
    n = 2000
    rng = np.random.default_rng(4)
    a = tri(rng, n, True)
    b = rng.standard_normal(n).astype(np.float32)
    x, _ = L.trsv(a, b, True, False, dtype=np.float32)
    xo = oracle.solve_triangular(a, b, True, False)
    exact = oracle.solve_triangular(a.astype(np.float64), b.astype(np.float64), True, False)
    scale = np.abs(exact).max()
    err_gpu = np.abs(x.astype(np.float64) - exact).max() / scale
    err_orc = np.abs(xo.astype(np.float64) - exact).max() / scale
    # both f32 loops drift several ulps from the exact solution; an f64 solve
    # rounded to f32 would be within half an ulp (2^-24) of it
    assert err_orc > 4 * U32
    assert err_gpu > 2 * U32, err_gpu
    assert err_gpu <= 8 * n * U32 * np.linalg.cond(a.astype(np.float64), np.inf)


@pytest.mark.parametrize("n,nrhs", [(5, 1), (64, 2), (300, 1), (1000, 4)])
def test_sposv_matches_oracle_f32(L, n, nrhs):
    rng = np.random.default_rng(7 * n + nrhs)
    g = rng.standard_normal((n, n))
    s = (g @ g.T / n + np.eye(n)).astype(np.float32)
    b = rng.standard_normal((n, nrhs)).astype(np.float32)
    print(f"sposv n={n} nrhs={nrhs}", flush=True)
    x, _ = L.posv(s, b, dtype=np.float32)
    assert x.dtype == np.float32
    xo = oracle.cholesky_solve(s, b)
    assert xo.dtype == np.float32
    kappa = np.linalg.cond(s.astype(np.float64), np.inf)
    err = np.abs(x.astype(np.float64) - xo).max() / np.abs(xo).max()
    assert err <= 8 * n * U32 * kappa, err
    assert bwd32(s, x, b) <= 8 * n * U32


def test_sposv_not_pd_row_matches_oracle(L):
    rng = np.random.default_rng(5)
    g = rng.standard_normal((260, 260))
    a = (g @ g.T / 260 + np.eye(260)).astype(np.float32)
    a[200, 200] = -1.0
    with pytest.raises(oracle.OracleError) as eo:
        oracle.cholesky(a)
    with pytest.raises(L.NotPositiveDefiniteError) as eg:
        L.posv(a, np.ones(260), dtype=np.float32)
    assert eg.value.rc == eo.value.code == 201
