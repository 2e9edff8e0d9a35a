"""Batched uneven solves against the oracle, system by system.

Each system is solved by the reference's per-system semantics (solve.rs:54-111
for general systems, cholesky + two substitutions for SPD ones); the batched
path factors in FP32 and refines in FP64, so the bar is the FP64 one:
normwise backward error <= 8 * n * 2^-53 per right-hand side, and forward
error against the oracle within 8 * n * 2^-53 * kappa_inf.  Breakdowns must be
reported at the oracle's column (LU) / row (Cholesky).
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


def make_batch(count, nmin, nmax, seed, kappa=1e3):
    rng = np.random.default_rng(seed)
    As, Bs, spd = [], [], []
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        nrhs = 1 if rng.random() < 0.5 else 4
        if i % 2 == 0:
            u, _ = np.linalg.qr(rng.standard_normal((n, n)))
            v, _ = np.linalg.qr(rng.standard_normal((n, n)))
            a = (u * np.logspace(0, np.log10(kappa), n)) @ v.T
            spd.append(0)
        else:
            g = rng.standard_normal((n, n))
            a = g @ g.T / n + np.eye(n)
            a = 0.5 * (a + a.T)
            spd.append(1)
        As.append(a)
        Bs.append(rng.standard_normal((n, nrhs)))
    return As, Bs, np.array(spd, dtype=np.uint8)


def bwd(a, x, b):
    r = b - a @ x
    na = np.abs(a).sum(axis=1).max()
    return max(np.abs(r[:, j]).max() / (na * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(x.shape[1]))


@pytest.mark.parametrize("nmin,nmax,count,seed", [(8, 256, 120, 1), (8, 48, 200, 2), (200, 256, 24, 3)])
def test_batched_matches_oracle(L, nmin, nmax, count, seed):
    As, Bs, spd = make_batch(count, nmin, nmax, seed)
    xs, status, eta, info = L.gesv_batched(As, Bs, spd)
    assert info.failed == 0, status
    for i, (a, b, x) in enumerate(zip(As, Bs, xs)):
        n = a.shape[0]
        e = bwd(a, x, b)
        assert e <= 8 * n * U64, (i, n, e)
        assert eta[i] <= 8 * n * U64
        if n <= 160 and i % 3 == 0:
            xo = oracle.cholesky_solve(a, b) if spd[i] else oracle.solve_multiple(a, b)
            kappa = np.linalg.cond(a, np.inf)
            assert np.abs(x - xo).max() / np.abs(xo).max() <= 8 * n * U64 * kappa, (i, n)


def test_batched_breakdowns_match_oracle(L):
    rng = np.random.default_rng(9)
    As, Bs, spd = make_batch(16, 20, 90, 4)
    # a general system with an exactly zero column -> singular at that column
    a = rng.standard_normal((40, 40)); a[:, 17] = 0.0
    As.append(a); Bs.append(rng.standard_normal((40, 1))); spd = np.append(spd, 0)
    # an SPD-flagged system that is not positive definite
    g = rng.standard_normal((50, 50)); s = g @ g.T / 50 + np.eye(50); s[30, 30] = -2.0
    As.append(s); Bs.append(rng.standard_normal((50, 1))); spd = np.append(spd, 1)
    xs, status, eta, info = L.gesv_batched(As, Bs, spd)
    with pytest.raises(oracle.OracleError) as e1:
        oracle.lu_factor(As[-2])
    with pytest.raises(oracle.OracleError) as e2:
        oracle.cholesky(As[-1])
    assert status[-2] == e1.value.code == 18
    assert status[-1] == e2.value.code
    assert info.failed == 2
    assert np.all(status[:-2] == 0)


def test_batched_graph_replay_is_identical(L):
    import torch
    As, Bs, spd = make_batch(64, 8, 256, 5)
    n, nrhs, offA, offB, A, B = L.pack_batch(As, Bs)
    dA = torch.from_numpy(A).cuda(); dB = torch.from_numpy(B).cuda()
    dX1 = torch.zeros_like(dB)
    st1, e1, i1 = L.gesv_batched_dev(n, nrhs, offA, offB, dA, dB, dX1, spd)
    x1 = dX1.clone()
    # same plan, same buffers: the captured graph is replayed
    torch.cuda.synchronize()
    st2, e2, i2 = L.gesv_batched_dev(n, nrhs, offA, offB, dA, dB, dX1, spd)
    st3, e3, i3 = L.gesv_batched_dev(n, nrhs, offA, offB, dA, dB, dX1, spd)
    assert i2.graph_replayed == 1 and i3.graph_replayed == 1
    assert np.array_equal(st1, st2) and np.array_equal(st2, st3)
    assert torch.equal(x1, dX1)
    assert i1.failed == 0


def test_host_entry_point_matches_device_entry(L):
    """scirs_gpu_dgesv_batched (host pointers, per-system strides) gives the
    device entry point's solutions bit for bit: the same packed systems reach
    the same kernels.  Strided views (lda > n) and a repeated call (graph
    replay on the kept buffers) are covered."""
    As, Bs, spd = make_batch(60, 8, 200, 11)
    wide = [np.pad(a, ((0, 0), (0, 5)))[:, :a.shape[0]] for a in As]   # row stride n + 5
    for a, w in zip(As, wide):
        assert np.array_equal(a, w) and w.strides[0] == 8 * (a.shape[0] + 5)
    xs_d, st_d, eta_d, _ = L.gesv_batched(As, Bs, spd)
    for call in range(2):
        xs_h, st_h, eta_h, info = L.gesv_batched_host(wide, Bs, spd)
        assert info.failed == 0 and np.array_equal(st_h, st_d)
        for i, (a, b) in enumerate(zip(As, Bs)):
            assert np.array_equal(xs_h[i], xs_d[i]), (call, i)
            assert bwd(a, xs_h[i], b) <= 8 * a.shape[0] * U64
        assert info.total_ms >= info.time_ms > 0


def test_host_entry_point_rejects_bad_strides(L):
    a = np.eye(4)
    with pytest.raises(ValueError):
        L.gesv_batched_host([a[:, :3]], [np.ones(4)])   # not square
    import ctypes
    from scirs_amd import _native as N
    lib = N.load()
    n = np.array([4], np.int32); r = np.array([1], np.int32)
    lda = np.array([3], np.int64); ld1 = np.array([1], np.int64)
    A = (ctypes.c_void_p * 1)(a.ctypes.data); B = (ctypes.c_void_p * 1)(a.ctypes.data)
    st = np.zeros(1, np.int32)
    rc = lib.scirs_gpu_dgesv_batched(1, n.ctypes.data, r.ctypes.data, None, ctypes.cast(A, ctypes.c_void_p),
                                     lda.ctypes.data, ctypes.cast(B, ctypes.c_void_p), ld1.ctypes.data,
                                     ctypes.cast(B, ctypes.c_void_p), ld1.ctypes.data, st.ctypes.data, None, 0, 0.0,
                                     None)
    assert rc == N.SCIRS_ERR_ARG and "lda=3" in N.last_error()
