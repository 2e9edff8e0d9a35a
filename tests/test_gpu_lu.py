"""Parity of the HIP LU / Cholesky path against the oracle (through the C ABI).

Bars (floating point; tolerances are written in each test):
* normwise backward error  eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)
  <= 8 * n * 2^-53 (FP64) — the c*n*u bound of the north star with c = 8;
* forward error against the oracle's x within 8 * n * u * kappa_inf;
* pivot sequence equal to the oracle's on the golden inputs (same first-max
  rule); the breakdown column of singular inputs equal to the oracle's.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
U64 = 2.0 ** -53
U32 = 2.0 ** -24


@pytest.fixture(scope="module")
def L():
    from scirs_amd import linalg
    return linalg


def bwd(a, x, b):
    a = np.asarray(a, np.float64)
    x = np.asarray(x, np.float64).reshape(a.shape[0], -1)
    b = np.asarray(b, np.float64).reshape(a.shape[0], -1)
    r = b - a @ x
    na = np.abs(a).sum(axis=1).max()
    return max(np.abs(r[:, j]).max() / (na * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(x.shape[1]))


def gold(n):
    return np.load(os.path.join(GOLD, f"lapack_n{n}.npz"))


with open(os.path.join(GOLD, "reference_kats.json")) as f:
    KATS = json.load(f)


def test_lu_kat_exact(L):
    k = KATS["lu_factor_2x2"]
    lu, perm, _ = L.getrf(np.array(k["a"]))
    assert np.array_equal(lu, np.array(k["lu"]))
    assert list(perm) == k["piv"]


@pytest.mark.parametrize("n", [5, 8, 16, 64, 512])
def test_getrf_matches_oracle(L, n):
    d = gold(n)
    lu, perm, _ = L.getrf(d["a"])
    lu_o, piv_o, _ = oracle.lu_factor(d["a"])
    assert np.array_equal(perm, piv_o), "pivot sequence differs from the oracle"
    assert np.array_equal(perm, d["perm"])
    # factors agree to rounding (blocked vs unblocked summation order)
    assert np.abs(lu - lu_o).max() <= 64 * n * U64 * max(1.0, np.abs(lu_o).max())
    l = np.tril(lu, -1) + np.eye(n)
    u = np.triu(lu)
    res = np.abs(d["a"][perm] - l @ u).max() / np.abs(d["a"]).max()
    assert res <= 8 * n * U64


@pytest.mark.parametrize("n", [5, 8, 16, 64, 512])
def test_gesv_matches_oracle(L, n):
    d = gold(n)
    x, info = L.gesv(d["a"], d["b"])
    xo = oracle.solve(d["a"], d["b"])
    assert bwd(d["a"], x, d["b"]) <= 8 * n * U64
    kappa = np.linalg.cond(d["a"], np.inf)
    assert np.abs(x - xo).max() / np.abs(xo).max() <= 8 * n * U64 * kappa
    xm, _ = L.gesv(d["a"], d["bm"])
    xmo = oracle.solve_multiple(d["a"], d["bm"])
    assert bwd(d["a"], xm, d["bm"]) <= 8 * n * U64
    assert np.abs(xm - xmo).max() / np.abs(xmo).max() <= 8 * n * U64 * kappa


@pytest.mark.parametrize("n,nrhs,seed", [(1000, 1, 1), (1536, 3, 2), (4096, 16, 3)])
def test_gesv_large_backward_error(L, n, nrhs, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    b = rng.standard_normal((n, nrhs))
    x, info = L.gesv(a, b)
    assert bwd(a, x, b) <= 8 * n * U64
    if n <= 1536:  # oracle cross-check where it finishes in seconds
        xo = oracle.solve_multiple(a, b)
        kappa = np.linalg.cond(a, np.inf)
        assert np.abs(x - xo).max() / np.abs(xo).max() <= 8 * n * U64 * kappa
        _, perm, _ = L.getrf(a)
        _, piv_o, _ = oracle.lu_factor(a)
        assert np.array_equal(perm, piv_o)


def test_getrf_rectangular(L):
    rng = np.random.default_rng(9)
    for shape in ((300, 200), (200, 300), (130, 129)):
        a = rng.standard_normal(shape)
        lu, perm, _ = L.getrf(a)
        p_o, l_o, u_o = oracle.lu(a)
        m, n = shape
        k = min(m, n)
        l = np.tril(lu, -1)[:, :k] + np.eye(m, k)
        u = np.triu(lu)[:k, :]
        assert np.array_equal(perm, np.argmax(p_o, axis=1))
        assert np.abs(a[perm] - l @ u).max() <= 8 * max(m, n) * U64 * np.abs(a).max()


def test_singular_column_matches_oracle(L):
    rng = np.random.default_rng(4)
    # a zero column stays exactly zero through the elimination: singular at its column
    a = rng.standard_normal((300, 300))
    a[:, 150] = 0.0
    with pytest.raises(oracle.OracleError) as eo:
        oracle.lu_factor(a)
    with pytest.raises(L.SingularError) as eg:
        L.getrf(a)
    assert eg.value.rc == eo.value.code == 151
    # a zero row is never a pivot until the last column
    a = rng.standard_normal((300, 300))
    a[77, :] = 0.0
    with pytest.raises(oracle.OracleError) as eo:
        oracle.lu_factor(a)
    with pytest.raises(L.SingularError) as eg:
        L.getrf(a)
    assert eg.value.rc == eo.value.code == 300
    # absolute eps rule: a well-conditioned matrix scaled below eps is singular at column 1
    with pytest.raises(L.SingularError) as eg:
        L.getrf(np.eye(64) * 1e-17)
    assert eg.value.rc == 1
    with pytest.raises(L.SingularError) as eg:
        L.gesv(np.zeros((10, 10)), np.ones(10))
    assert eg.value.rc == 1


def test_singular_after_larger_factorisation(L):
    """Regression: a factorisation that stops at a singular column must not pick
    up the row-interchange lists an earlier, larger factorisation left for the
    panels it never ran (they named rows beyond n)."""
    rng = np.random.default_rng(21)
    big = rng.standard_normal((777, 777))
    L.gesv(big, rng.standard_normal(777))
    a = rng.standard_normal((200, 200))
    a[:, 57] = 0.0
    with pytest.raises(oracle.OracleError) as eo:
        oracle.lu_factor(a)
    for call in (lambda: L.gesv(a, np.ones(200)), lambda: L.getrf(a)):
        with pytest.raises(L.SingularError) as eg:
            call()
        assert eg.value.rc == eo.value.code == 58
    # the library is still usable afterwards
    x, _ = L.gesv(big[:100, :100], np.ones(100))
    assert np.all(np.isfinite(x))


def test_tie_rule_first_max(L):
    a = np.array([[1.0, 2.0, 0.0, 1.0, 0.0],
                  [0.5, 1.0, 3.0, 0.0, 1.0],
                  [1.0, 0.0, 1.0, 2.0, 2.0],
                  [-1.0, 1.0, 1.0, 1.0, 3.0],
                  [0.25, 1.0, 2.0, 1.0, 1.0]])
    _, perm, _ = L.getrf(a)
    _, piv_o, _ = oracle.lu_factor(a)
    assert np.array_equal(perm, piv_o)


@pytest.mark.parametrize("n,dtype", [(1024, np.float64), (2000, np.float64), (1536, np.float32)])
def test_getrf_exchange_path_matches_oracle(L, n, dtype):
    """Sizes whose panels span several workgroups (the tagged candidate
    exchange, the fused next-block kernel and the look-ahead streams; the
    golden sizes above all fit one workgroup per panel): the pivot sequence
    equals the oracle's and P A = L U to rounding."""
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)).astype(dtype)
    lu, perm, _ = L.getrf(a)
    _, piv_o, _ = oracle.lu_factor(a)
    assert np.array_equal(perm, piv_o), "pivot sequence differs from the oracle"
    u_ = U64 if dtype == np.float64 else U32
    a64 = a.astype(np.float64)
    lu = lu.astype(np.float64)
    l = np.tril(lu, -1) + np.eye(n)
    u = np.triu(lu)
    res = np.abs(a64[perm] - l @ u).max() / np.abs(a64).max()
    assert res <= 8 * n * u_


def test_sgetrf_sgesv(L):
    d = gold(512)
    a32 = d["a"].astype(np.float32)
    b32 = d["b"].astype(np.float32)
    x, _ = L.gesv(a32, b32)
    assert x.dtype == np.float32
    assert bwd(a32.astype(np.float64), x, b32) <= 8 * 512 * U32
    _, perm, _ = L.getrf(a32)
    _, piv_o, _ = oracle.lu_factor(a32)
    assert np.array_equal(perm, piv_o)


# ------------------------------------------------------------------ Cholesky
def test_cholesky_kats_exact(L):
    for key in ("cholesky_2x2", "cholesky_3x3_textbook"):
        k = KATS[key]
        l, _ = L.potrf(np.array(k["a"]))
        assert np.array_equal(l, np.array(k["l"]))
    for m in KATS["cholesky_spd_set"]["mats"]:
        a = np.array(m)
        l, _ = L.potrf(a)
        assert np.allclose(l, oracle.cholesky(a), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("n", [5, 8, 16, 64])
def test_potrf_matches_oracle(L, n):
    d = gold(n)
    l, _ = L.potrf(d["spd"])
    lo = oracle.cholesky(d["spd"])
    assert np.all(np.triu(l, 1) == 0)
    assert np.abs(l - lo).max() <= 32 * n * U64 * np.abs(lo).max()
    x, _ = L.posv(d["spd"], d["b"])
    xo = oracle.cholesky_solve(d["spd"], d["b"])
    assert bwd(d["spd"], x, d["b"]) <= 8 * n * U64
    kappa = np.linalg.cond(d["spd"], np.inf)
    assert np.abs(x - xo).max() / np.abs(xo).max() <= 8 * n * U64 * kappa


@pytest.mark.parametrize("n,nrhs", [(300, 1), (1000, 4), (2048, 16)])
def test_posv_large(L, n, nrhs):
    rng = np.random.default_rng(n)
    g = rng.standard_normal((n, n))
    a = g @ g.T / n + np.eye(n)
    b = rng.standard_normal((n, nrhs))
    x, _ = L.posv(a, b)
    assert bwd(a, x, b) <= 8 * n * U64
    l, _ = L.potrf(a)
    if n <= 1000:
        lo = oracle.cholesky(a)
        assert np.abs(l - lo).max() <= 32 * n * U64 * np.abs(lo).max()


def test_potrf_reads_lower_only(L):
    rng = np.random.default_rng(3)
    g = rng.standard_normal((200, 200))
    a = g @ g.T / 200 + np.eye(200)
    junk = a.copy()
    junk[np.triu_indices(200, 1)] = 1e300
    l1, _ = L.potrf(a)
    l2, _ = L.potrf(junk)
    assert np.array_equal(l1, l2)


def test_potrf_not_pd_row_matches_oracle(L):
    rng = np.random.default_rng(5)
    g = rng.standard_normal((260, 260))
    a = g @ g.T / 260 + np.eye(260)
    a[200, 200] = -1.0
    with pytest.raises(oracle.OracleError) as eo:
        oracle.cholesky(a)
    with pytest.raises(L.NotPositiveDefiniteError) as eg:
        L.potrf(a)
    assert eg.value.rc == eo.value.code == 201


# ------------------------------------------------------------------ triangular
def test_trsv_kats(L):
    for c in KATS["solve_triangular_cases"]["cases"]:
        x, _ = L.trsv(np.array(c["a"]), np.array(c["b"]), c["lower"], c["unit"])
        assert np.allclose(x, c["x"], rtol=1e-15, atol=0)


@pytest.mark.parametrize("lower,unit", [(True, True), (True, False), (False, True), (False, False)])
def test_trsv_matches_oracle(L, lower, unit):
    rng = np.random.default_rng(7)
    n = 333
    a = rng.standard_normal((n, n)) / np.sqrt(n)
    a = np.tril(a) if lower else np.triu(a)
    np.fill_diagonal(a, 1.0 + rng.random(n))
    b = rng.standard_normal(n)
    x, _ = L.trsv(a, b, lower, unit)
    xo = oracle.solve_triangular(a, b, lower, unit)
    assert np.abs(x - xo).max() / np.abs(xo).max() <= 1e-12


@pytest.mark.parametrize("n", [1, 63, 64, 65, 200, 1000])
@pytest.mark.parametrize("nrhs", [1, 2, 3, 7])
def test_trsv_chain_sizes_and_rhs(L, n, nrhs):
    """The one-launch chained solver: partial last block row (n % 64 != 0),
    right-hand sides split into chunks of 2 (FP64), every triangle kind.
    Bar: per column, |x - x_oracle| <= 1e-12 max|x_oracle| (well-conditioned
    triangles: unit-ish diagonal, off-diagonal entries ~ 1/sqrt(n))."""
    rng = np.random.default_rng(1000 * n + nrhs)
    for lower in (True, False):
        for unit in (True, False):
            a = rng.standard_normal((n, n)) / np.sqrt(n)
            a = np.tril(a) if lower else np.triu(a)
            np.fill_diagonal(a, 1.0 + rng.random(n))
            b = rng.standard_normal((n, nrhs))
            print(f"trsv n={n} nrhs={nrhs} lower={lower} unit={unit}", flush=True)
            x, _ = L.trsv(a, b, lower, unit)
            for j in range(nrhs):
                xo = oracle.solve_triangular(a, np.ascontiguousarray(b[:, j]), lower, unit)
                assert np.abs(x[:, j] - xo).max() <= 1e-12 * np.abs(xo).max(), (lower, unit, j)


def test_trsv_singular_diagonal(L):
    a = np.array([[1.0, 0.0, 0.0], [2.0, 1e-20, 0.0], [1.0, 1.0, 1.0]])
    with pytest.raises(L.SingularError) as eg:
        L.trsv(a, np.ones(3), True, False)
    assert eg.value.rc == 2


def test_dev_entry_points_follow_callers_default_stream(L):
    """Inputs produced by torch on the default stream and handed over at once
    (no synchronisation): the device entry points must queue behind that work.
    Regression for a factorisation that read a half-written matrix."""
    import torch
    n = 6144
    g = torch.Generator(device="cuda").manual_seed(7)
    gg = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    s = gg @ gg.T.contiguous() / n + torch.eye(n, dtype=torch.float64, device="cuda")
    b = torch.randn(n, 2, dtype=torch.float64, device="cuda", generator=g)
    x = torch.empty_like(b)
    L.posv_dev(s, b, x, torch.empty_like(s))
    assert L.backward_error_dev(s, x, b) <= 8 * n * 2.0 ** -53
    a = gg @ gg.T.contiguous() / n + torch.eye(n, dtype=torch.float64, device="cuda")
    x2 = torch.empty_like(b)
    L.gesv_dev(a, b, x2, torch.empty_like(a), torch.empty(n, dtype=torch.int64, device="cuda"))
    assert L.backward_error_dev(a, x2, b) <= 8 * n * 2.0 ** -53
