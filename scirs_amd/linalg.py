"""numpy / torch front end of the C ABI (test and benchmark harness).

Host arrays go through the host-memory entry points; torch CUDA tensors go
through the *_dev entry points without leaving HBM.  Every routine calls the
HIP library; there is no CPU path here.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N


class GpuSolveError(RuntimeError):
    def __init__(self, fn: str, rc: int, info: N.SolveInfo | None = None):
        msg = N.last_error() if rc < 0 else ""
        super().__init__(f"{fn} returned {rc} {msg}".strip())
        self.rc = rc
        self.info = info


class SingularError(GpuSolveError):
    """rc > 0 from an LU routine: |pivot| < eps at 1-based column rc."""


class NotPositiveDefiniteError(GpuSolveError):
    """rc > 0 from a Cholesky routine: a_ii - sum <= 0 at 1-based row rc."""


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _check(fn, rc, info, chol=False):
    if rc > 0:
        raise (NotPositiveDefiniteError if chol else SingularError)(fn, rc, info)
    if rc < 0:
        raise GpuSolveError(fn, rc, info)


def getrf(a):
    """Packed LU and the reference-style permutation vector (P A = L U)."""
    lib = N.load()
    a = np.asarray(a)
    dt = np.float32 if a.dtype == np.float32 else np.float64
    lu = _c(a, dt).copy()
    m, n = lu.shape
    perm = np.empty(m, dtype=np.int64)
    info = N.SolveInfo()
    fn = lib.scirs_gpu_sgetrf if dt == np.float32 else lib.scirs_gpu_dgetrf
    rc = fn(m, n, _ptr(lu), n, _ptr(perm), ctypes.byref(info))
    _check("getrf", rc, info)
    return lu, perm, info


def gesv(a, b):
    lib = N.load()
    a = np.asarray(a)
    dt = np.float32 if a.dtype == np.float32 else np.float64
    a = _c(a, dt)
    b = np.asarray(b, dtype=dt)
    vec = b.ndim == 1
    b2 = _c(b.reshape(b.shape[0], -1), dt)
    n, nrhs = b2.shape
    x = np.empty_like(b2)
    info = N.SolveInfo()
    fn = lib.scirs_gpu_sgesv if dt == np.float32 else lib.scirs_gpu_dgesv
    rc = fn(n, nrhs, _ptr(a), a.shape[1], _ptr(b2), nrhs, _ptr(x), nrhs, ctypes.byref(info))
    _check("gesv", rc, info)
    return (x[:, 0] if vec else x), info


def potrf(a):
    lib = N.load()
    a = np.asarray(a)
    dt = np.float32 if a.dtype == np.float32 else np.float64
    l = _c(a, dt).copy()
    n = l.shape[0]
    info = N.SolveInfo()
    fn = lib.scirs_gpu_spotrf if dt == np.float32 else lib.scirs_gpu_dpotrf
    rc = fn(n, _ptr(l), n, ctypes.byref(info))
    _check("potrf", rc, info, chol=True)
    return l, info


def posv(a, b, dtype=np.float64):
    """Cholesky solve in f64 (scirs_gpu_dposv) or f32 (scirs_gpu_sposv)."""
    lib = N.load()
    dt = np.dtype(dtype)
    a = _c(a, dt)
    b = np.asarray(b, dtype=dt)
    vec = b.ndim == 1
    b2 = _c(b.reshape(b.shape[0], -1), dt)
    n, nrhs = b2.shape
    x = np.empty_like(b2)
    info = N.SolveInfo()
    fn = lib.scirs_gpu_sposv if dt == np.float32 else lib.scirs_gpu_dposv
    rc = fn(n, nrhs, _ptr(a), n, _ptr(b2), nrhs, _ptr(x), nrhs, ctypes.byref(info))
    _check("posv", rc, info, chol=True)
    return (x[:, 0] if vec else x), info


def trsv(a, b, lower, unit_diagonal, dtype=np.float64):
    """Triangular solve in f64 (scirs_gpu_dtrsv) or f32 (scirs_gpu_strsv)."""
    lib = N.load()
    dt = np.dtype(dtype)
    a = _c(a, dt)
    b = np.asarray(b, dtype=dt)
    vec = b.ndim == 1
    b2 = _c(b.reshape(b.shape[0], -1), dt)
    n, nrhs = b2.shape
    x = np.empty_like(b2)
    info = N.SolveInfo()
    fn = lib.scirs_gpu_strsv if dt == np.float32 else lib.scirs_gpu_dtrsv
    rc = fn(n, nrhs, _ptr(a), n, _ptr(b2), nrhs, _ptr(x), nrhs, int(lower), int(unit_diagonal),
            ctypes.byref(info))
    _check("trsv", rc, info)
    return (x[:, 0] if vec else x), info


def qr(a):
    """(Q, R) of an f64 rows x cols matrix, rows >= cols (lapack.rs:171-253)."""
    lib = N.load()
    a = _c(a, np.float64)
    rows, cols = a.shape
    q = np.empty((rows, rows))
    r = np.empty((rows, cols))
    info = N.SolveInfo()
    rc = lib.scirs_gpu_dqr(rows, cols, _ptr(a), cols, _ptr(q), rows, _ptr(r), cols, ctypes.byref(info))
    _check("qr", rc, info)
    return q, r


def lstsq(a, b):
    """(x, residuals) of the rows >= cols least-squares solve (solve.rs:234-290)."""
    lib = N.load()
    a = _c(a, np.float64)
    b = _c(b, np.float64)
    rows, cols = a.shape
    x = np.empty(cols)
    res = ctypes.c_double(0.0)
    info = N.SolveInfo()
    rc = lib.scirs_gpu_dlstsq(rows, cols, _ptr(a), cols, _ptr(b), _ptr(x), ctypes.byref(res), ctypes.byref(info))
    _check("lstsq", rc, info)
    return x, res.value


def gmres(a, b, x0=None, max_iter=1000, tol=1e-10, restart=0):
    """Restarted GMRES as solvers/iterative.rs:317-450: (x, iterations, residual_norm, converged)."""
    lib = N.load()
    a = _c(a, np.float64)
    b = _c(b, np.float64)
    n = a.shape[0]
    x = np.empty(n)
    it = ctypes.c_int(0)
    conv = ctypes.c_int(0)
    res = ctypes.c_double(0.0)
    x0p = None if x0 is None else _c(x0, np.float64)
    info = N.SolveInfo()
    rc = lib.scirs_gpu_dgmres(n, _ptr(a), n, _ptr(b), None if x0p is None else _ptr(x0p), int(max_iter),
                              float(tol), int(restart), _ptr(x), ctypes.byref(it), ctypes.byref(res),
                              ctypes.byref(conv), ctypes.byref(info))
    _check("gmres", rc, info)
    return x, it.value, res.value, bool(conv.value)


def gmres_precond(a, b, precond, lu=None, perm=None, x0=None, max_iter=1000, tol=1e-10, restart=0):
    """Right-preconditioned restarted GMRES (the real counterpart of
    preconditioners.rs:934-1010): precond is N.PRECOND_NONE / _DIAGONAL / _LU /
    _LU_GIVEN (lu: packed FP64 factors, perm: P A = L U).  Returns
    (x, iterations, residual_norm, converged)."""
    lib = N.load()
    a = _c(a, np.float64)
    b = _c(b, np.float64)
    n = a.shape[0]
    x = np.empty(n)
    it = ctypes.c_int(0)
    conv = ctypes.c_int(0)
    res = ctypes.c_double(0.0)
    x0p = None if x0 is None else _c(x0, np.float64)
    lup = None if lu is None else _c(lu, np.float64)
    pp = None if perm is None else np.ascontiguousarray(perm, dtype=np.int64)
    info = N.SolveInfo()
    rc = lib.scirs_gpu_dgmres_precond(n, _ptr(a), n, _ptr(b), None if x0p is None else _ptr(x0p), int(max_iter),
                                      float(tol), int(restart), int(precond), None if lup is None else _ptr(lup),
                                      None if pp is None else _ptr(pp), _ptr(x), ctypes.byref(it),
                                      ctypes.byref(res), ctypes.byref(conv), ctypes.byref(info))
    _check("gmres_precond", rc, info)
    return x, it.value, res.value, bool(conv.value)


def gmres_dev(a, b, x, use_x0=False, max_iter=1000, tol=1e-10, restart=0):
    """gmres on device tensors (float64, cuda): x holds x0 when use_x0 and the
    solution on return.  Returns (iterations, residual_norm, converged)."""
    lib = N.load()
    n = a.shape[0] if hasattr(a, "shape") and len(a.shape) == 2 else -1
    if n <= 0 or a.shape[1] != n:
        # the reference's shape text (solvers/iterative.rs:327-333)
        raise ValueError(f"gmres_dev: Expected square matrix, got shape {tuple(getattr(a, 'shape', ()))}")
    _dev_arg("gmres_dev", "a", a, n, n)
    _dev_arg("gmres_dev", "b", b, n)
    _dev_arg("gmres_dev", "x", x, n)
    it = ctypes.c_int(0)
    conv = ctypes.c_int(0)
    res = ctypes.c_double(0.0)
    info = N.SolveInfo()
    rc = lib.scirs_gpu_dgmres_dev(n, ctypes.c_void_p(a.data_ptr()), a.stride(0), ctypes.c_void_p(b.data_ptr()),
                                  ctypes.c_void_p(x.data_ptr()), int(use_x0), int(max_iter), float(tol),
                                  int(restart), ctypes.byref(it), ctypes.byref(res), ctypes.byref(conv),
                                  _stream_ptr(), ctypes.byref(info))
    _check("gmres_dev", rc, info)
    return it.value, res.value, bool(conv.value)


def _batch_args(a):
    a = np.asarray(a)
    dt = np.float32 if a.dtype == np.float32 else np.float64
    return _c(a, dt), dt, ("s" if dt == np.float32 else "d")


def batch_matmul(a, b):
    """(batch, m, k) x (k, n) -> (batch, m, n), bit-identical to batch/mod.rs:62-107."""
    lib = N.load()
    a, dt, p = _batch_args(a)
    b = _c(b, dt)
    nb, m, k = a.shape
    n = b.shape[1]
    c = np.empty((nb, m, n), dtype=dt)
    info = N.SolveInfo()
    rc = getattr(lib, f"scirs_gpu_{p}batch_matmul")(nb, m, k, n, _ptr(a), _ptr(b), _ptr(c), ctypes.byref(info))
    _check("batch_matmul", rc, info)
    return c


def batch_matvec(a, x):
    """(batch, m, n) x (n,) -> (batch, m), bit-identical to batch/mod.rs:139-172."""
    lib = N.load()
    a, dt, p = _batch_args(a)
    x = _c(x, dt)
    nb, m, n = a.shape
    y = np.empty((nb, m), dtype=dt)
    info = N.SolveInfo()
    rc = getattr(lib, f"scirs_gpu_{p}batch_matvec")(nb, m, n, _ptr(a), _ptr(x), _ptr(y), ctypes.byref(info))
    _check("batch_matvec", rc, info)
    return y


def batch_add(a, v, axis):
    """a + v along axis 0 (rows) or 1 (columns), as batch/mod.rs:220-291."""
    lib = N.load()
    a, dt, p = _batch_args(a)
    v = _c(v, dt)
    nb, m, n = a.shape
    out = np.empty_like(a)
    info = N.SolveInfo()
    rc = getattr(lib, f"scirs_gpu_{p}batch_add")(nb, m, n, _ptr(a), _ptr(v), int(axis), _ptr(out),
                                                 ctypes.byref(info))
    _check("batch_add", rc, info)
    return out


def batch_sum(a):
    """The sum over the batch axis, in batch order (batch/mod.rs:326-336)."""
    lib = N.load()
    a, dt, p = _batch_args(a)
    nb, m, n = a.shape
    out = np.empty((m, n), dtype=dt)
    info = N.SolveInfo()
    rc = getattr(lib, f"scirs_gpu_{p}batch_sum")(nb, m, n, _ptr(a), _ptr(out), ctypes.byref(info))
    _check("batch_sum", rc, info)
    return out


def gesv_ir(a, b, mode=N.MODE_AUTO, max_iter=0, max_gmres=0, tol=0.0):
    lib = N.load()
    a = _c(a, np.float64)
    b = np.asarray(b, dtype=np.float64)
    vec = b.ndim == 1
    b2 = _c(b.reshape(b.shape[0], -1), np.float64)
    n, nrhs = b2.shape
    x = np.empty_like(b2)
    info = N.SolveInfo()
    rc = lib.scirs_gpu_dgesv_ir(n, nrhs, _ptr(a), n, _ptr(b2), nrhs, _ptr(x), nrhs, int(mode), int(max_iter),
                                int(max_gmres), float(tol), ctypes.byref(info))
    _check("gesv_ir", rc, info)
    return (x[:, 0] if vec else x), info


# ---------------------------------------------------------------- torch (HBM)
def _tp(t):
    return ctypes.c_void_p(t.data_ptr())


def _dev_arg(fn, name, t, rows, cols=None, dtype=None):
    """A device operand as the C ABI reads it: a CUDA tensor of the given dtype
    (float64 unless stated), `rows` rows (and `cols` columns when 2-D) with
    unit column stride; a transposed or strided view would be read wrongly and
    a short one past its end, so they are refused here."""
    import torch
    dtype = torch.float64 if dtype is None else dtype
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{fn}: {name} must be a CUDA tensor")
    if t.dtype != dtype:
        raise ValueError(f"{fn}: {name} must be {dtype}, got {t.dtype}")
    if cols is None:
        if t.dim() != 1 or t.shape[0] != rows or (rows > 1 and t.stride(0) != 1):
            raise ValueError(f"{fn}: {name} must be a contiguous vector of length {rows}, got {tuple(t.shape)}")
    else:
        if t.dim() != 2 or t.shape[0] != rows or t.shape[1] != cols or (cols > 1 and t.stride(1) != 1) \
                or t.stride(0) < max(cols, 1):
            raise ValueError(f"{fn}: {name} must be {rows} x {cols} with unit column stride, "
                             f"got shape {tuple(t.shape)} strides {t.stride()}")
    return t


def _stream_ptr():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def gesv_dev(A, B, X, LU=None, perm=None):
    """A, B, X: contiguous float64 CUDA tensors (n x n, n x nrhs, n x nrhs)."""
    lib = N.load()
    n, nrhs = B.shape
    _dev_arg("gesv_dev", "A", A, n, n)
    _dev_arg("gesv_dev", "B", B, n, nrhs)
    _dev_arg("gesv_dev", "X", X, n, nrhs)
    if LU is not None:
        _dev_arg("gesv_dev", "LU", LU, n, n)
    if perm is not None:
        import torch
        _dev_arg("gesv_dev", "perm", perm, n, dtype=torch.int64)
    info = N.SolveInfo()
    rc = lib.scirs_gpu_dgesv_dev(n, nrhs, _tp(A), A.stride(0), _tp(B), B.stride(0), _tp(X), X.stride(0),
                                 _tp(LU) if LU is not None else None, LU.stride(0) if LU is not None else 0,
                                 _tp(perm) if perm is not None else None, _stream_ptr(), ctypes.byref(info))
    _check("gesv_dev", rc, info)
    return info


def posv_dev(A, B, X, L=None):
    lib = N.load()
    n, nrhs = B.shape
    _dev_arg("posv_dev", "A", A, n, n)
    _dev_arg("posv_dev", "B", B, n, nrhs)
    _dev_arg("posv_dev", "X", X, n, nrhs)
    if L is not None:
        _dev_arg("posv_dev", "L", L, n, n)
    info = N.SolveInfo()
    rc = lib.scirs_gpu_dposv_dev(n, nrhs, _tp(A), A.stride(0), _tp(B), B.stride(0), _tp(X), X.stride(0),
                                 _tp(L) if L is not None else None, L.stride(0) if L is not None else 0,
                                 _stream_ptr(), ctypes.byref(info))
    _check("posv_dev", rc, info, chol=True)
    return info


def gesv_ir_dev(A, B, X, mode=N.MODE_AUTO, max_iter=0, max_gmres=0, tol=0.0):
    lib = N.load()
    n, nrhs = B.shape
    _dev_arg("gesv_ir_dev", "A", A, n, n)
    _dev_arg("gesv_ir_dev", "B", B, n, nrhs)
    _dev_arg("gesv_ir_dev", "X", X, n, nrhs)
    info = N.SolveInfo()
    rc = lib.scirs_gpu_dgesv_ir_dev(n, nrhs, _tp(A), A.stride(0), _tp(B), B.stride(0), _tp(X), X.stride(0),
                                    int(mode), int(max_iter), int(max_gmres), float(tol), _stream_ptr(),
                                    ctypes.byref(info))
    _check("gesv_ir_dev", rc, info)
    return info


def backward_error_dev(A, X, B):
    lib = N.load()
    n, nrhs = B.shape
    _dev_arg("backward_error_dev", "A", A, n, n)
    _dev_arg("backward_error_dev", "X", X, n, nrhs)
    _dev_arg("backward_error_dev", "B", B, n, nrhs)
    out = ctypes.c_double(0.0)
    rc = lib.scirs_gpu_backward_error_dev(n, nrhs, _tp(A), A.stride(0), _tp(X), X.stride(0), _tp(B), B.stride(0),
                                          _stream_ptr(), ctypes.byref(out))
    _check("backward_error_dev", rc, None)
    return out.value


def backward_error_dev_noalloc(A, X, B):
    """backward_error_dev through the allocation-free entry point (fixed scratch;
    scirs_gpu_init must have run)."""
    lib = N.load()
    n, nrhs = B.shape
    _dev_arg("backward_error_dev_noalloc", "A", A, n, n)
    _dev_arg("backward_error_dev_noalloc", "X", X, n, nrhs)
    _dev_arg("backward_error_dev_noalloc", "B", B, n, nrhs)
    out = ctypes.c_double(0.0)
    rc = lib.scirs_gpu_backward_error_dev_noalloc(n, nrhs, _tp(A), A.stride(0), _tp(X), X.stride(0), _tp(B),
                                                  B.stride(0), _stream_ptr(), ctypes.byref(out))
    _check("backward_error_dev_noalloc", rc, None)
    return out.value


def sgetrf_tp_dev(A, perm, mode=N.MODE_FP32):
    """Tournament-pivoted FP32 LU in place on the n x n float32 CUDA tensor A
    (row stride a multiple of 4); perm (int64, n) receives the row permutation.
    Returns the status (0, or k > 0 for |u_kk| < eps)."""
    import torch
    n = A.shape[0]
    _dev_arg("sgetrf_tp_dev", "A", A, n, n, dtype=torch.float32)
    _dev_arg("sgetrf_tp_dev", "perm", perm, n, dtype=torch.int64)
    rc = N.load().scirs_gpu_sgetrf_tp_dev(n, _tp(A), A.stride(0), _tp(perm), int(mode), _stream_ptr())
    if rc < 0:
        _check("sgetrf_tp_dev", rc, None)
    return rc


def seq_plan(next_seq, span):
    """The exchange-tag arithmetic: (reset, base, next)."""
    base = ctypes.c_uint(0)
    nxt = ctypes.c_uint(0)
    reset = N.load().scirs_gpu_seq_plan(ctypes.c_uint(next_seq), ctypes.c_uint(span), ctypes.byref(base),
                                        ctypes.byref(nxt))
    return reset, base.value, nxt.value


def prof_begin():
    N.load().scirs_gpu_prof_begin()


def prof_end():
    """Per kernel class: launches, total_ms, flops, bytes (event timing)."""
    arr = (N.ProfStats * len(N.PROF_CLASSES))()
    rc = N.load().scirs_gpu_prof_end(arr)
    _check("prof_end", rc, None)
    return {name: {"launches": int(a.launches), "total_ms": a.total_ms, "flops": a.flops, "bytes": a.bytes}
            for name, a in zip(N.PROF_CLASSES, arr)}


# ---------------------------------------------------------------- batched
def pack_batch(As, Bs):
    """Host layout of a batch: packed FP64 buffers and per-system offsets."""
    n = np.array([a.shape[0] for a in As], dtype=np.int32)
    nrhs = np.array([b.reshape(b.shape[0], -1).shape[1] for b in Bs], dtype=np.int32)
    offA = np.zeros(len(As), dtype=np.int64)
    offB = np.zeros(len(As), dtype=np.int64)
    offA[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    offB[1:] = np.cumsum(n.astype(np.int64) * nrhs)[:-1]
    A = np.concatenate([np.ascontiguousarray(a, dtype=np.float64).ravel() for a in As])
    B = np.concatenate([np.ascontiguousarray(b, dtype=np.float64).reshape(b.shape[0], -1).ravel() for b in Bs])
    return n, nrhs, offA, offB, A, B


def gesv_batched_dev(n, nrhs, offA, offB, dA, dB, dX, spd=None, max_iter=0, tol_scale=0.0):
    """Batched solve on device buffers (torch CUDA tensors dA, dB, dX); the
    size / offset arrays are host numpy arrays.  Returns (status, bwd, info)."""
    lib = N.load()
    batch = len(n)
    n = np.ascontiguousarray(n, dtype=np.int32)
    nrhs = np.ascontiguousarray(nrhs, dtype=np.int32)
    offA = np.ascontiguousarray(offA, dtype=np.int64)
    offB = np.ascontiguousarray(offB, dtype=np.int64)
    spd_arr = None if spd is None else np.ascontiguousarray(spd, dtype=np.uint8)
    if not (len(nrhs) == len(offA) == len(offB) == batch) or (spd_arr is not None and len(spd_arr) != batch):
        raise ValueError("gesv_batched_dev: n, nrhs, offA, offB (and spd) must have one entry per system")
    need_a = int((offA + n.astype(np.int64) ** 2).max()) if batch else 0
    need_b = int((offB + n.astype(np.int64) * nrhs).max()) if batch else 0
    for name, t, need in (("dA", dA, need_a), ("dB", dB, need_b), ("dX", dX, need_b)):
        _dev_arg("gesv_batched_dev", name, t.reshape(-1) if t.is_contiguous() else t, t.numel())
        if t.numel() < need or not t.is_contiguous():
            raise ValueError(f"gesv_batched_dev: {name} must be contiguous with at least {need} elements")
    status = np.zeros(batch, dtype=np.int32)
    bwd = np.zeros(batch, dtype=np.float64)
    info = N.BatchInfo()
    rc = lib.scirs_gpu_dgesv_batched_dev(batch, _ptr(n), _ptr(nrhs), _ptr(spd_arr) if spd_arr is not None else None,
                                         _tp(dA), _ptr(offA), _tp(dB), _ptr(offB), _tp(dX), _ptr(status),
                                         _ptr(bwd), int(max_iter), float(tol_scale), _stream_ptr(),
                                         ctypes.byref(info))
    _check("gesv_batched_dev", rc, None)
    return status, bwd, info


def gesv_batched_host(As, Bs, spd=None, max_iter=0, tol_scale=0.0):
    """Batched solve through the host-memory entry point scirs_gpu_dgesv_batched:
    lists of f64 numpy systems (any row strides) in, list of solutions out.
    Returns (xs, status, bwd, info)."""
    lib = N.load()
    batch = len(As)
    if len(Bs) != batch or (spd is not None and len(spd) != batch):
        raise ValueError("gesv_batched_host: one B (and spd flag) per A")
    As = [np.asarray(a, dtype=np.float64) for a in As]
    Bs = [np.asarray(b, dtype=np.float64) for b in Bs]
    Bs = [b.reshape(-1, 1) if b.ndim == 1 else b for b in Bs]
    for i, (a, b) in enumerate(zip(As, Bs)):
        if a.ndim != 2 or a.shape[0] != a.shape[1] or b.ndim != 2 or b.shape[0] != a.shape[0]:
            raise ValueError(f"gesv_batched_host: system {i} has A {a.shape} and B {b.shape}")
    As = [a if a.strides[1] == 8 and a.strides[0] > 0 else np.ascontiguousarray(a) for a in As]
    Bs = [b if b.strides[1] == 8 and b.strides[0] > 0 else np.ascontiguousarray(b) for b in Bs]
    n = np.array([a.shape[0] for a in As], dtype=np.int32)
    nrhs = np.array([b.shape[1] for b in Bs], dtype=np.int32)
    xs = [np.zeros((int(n[i]), int(nrhs[i]))) for i in range(batch)]
    PA = (ctypes.c_void_p * batch)(*[a.ctypes.data for a in As])
    PB = (ctypes.c_void_p * batch)(*[b.ctypes.data for b in Bs])
    PX = (ctypes.c_void_p * batch)(*[x.ctypes.data for x in xs])
    lda = np.array([a.strides[0] // 8 for a in As], dtype=np.int64)
    ldb = np.array([b.strides[0] // 8 for b in Bs], dtype=np.int64)
    ldx = np.array([x.strides[0] // 8 for x in xs], dtype=np.int64)
    spd_arr = None if spd is None else np.ascontiguousarray(spd, dtype=np.uint8)
    status = np.zeros(batch, dtype=np.int32)
    bwd = np.zeros(batch, dtype=np.float64)
    info = N.BatchInfo()
    rc = lib.scirs_gpu_dgesv_batched(batch, _ptr(n), _ptr(nrhs), _ptr(spd_arr) if spd_arr is not None else None,
                                     ctypes.cast(PA, ctypes.c_void_p), _ptr(lda), ctypes.cast(PB, ctypes.c_void_p),
                                     _ptr(ldb), ctypes.cast(PX, ctypes.c_void_p), _ptr(ldx), _ptr(status),
                                     _ptr(bwd), int(max_iter), float(tol_scale), ctypes.byref(info))
    _check("gesv_batched", rc, None)
    return xs, status, bwd, info


def gesv_batched(As, Bs, spd=None, max_iter=0, tol_scale=0.0):
    """Host convenience wrapper: lists of numpy systems in, list of solutions out."""
    import torch
    n, nrhs, offA, offB, A, B = pack_batch(As, Bs)
    dA = torch.from_numpy(A).cuda()
    dB = torch.from_numpy(B).cuda()
    dX = torch.zeros_like(dB)
    status, bwd, info = gesv_batched_dev(n, nrhs, offA, offB, dA, dB, dX, spd, max_iter, tol_scale)
    X = dX.cpu().numpy()
    xs = [X[offB[i]:offB[i] + n[i] * nrhs[i]].reshape(n[i], nrhs[i]) for i in range(len(As))]
    return xs, status, bwd, info
