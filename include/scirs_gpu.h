/*
 * scirs_gpu.h — C ABI of the MI355X (gfx950) dense linear-solve engine.
 *
 * This is the whole boundary between a host binding (the scirs2-linalg Rust
 * shim, a ctypes test harness, ...) and the HIP kernels.  Plain pointers and
 * sizes only.  Matrices are dense and ROW-MAJOR (an ndarray in standard
 * order) with an explicit leading dimension in elements.
 *
 * Return value of every call (also copied to info->status):
 *      0   success
 *    k>0   the reference's breakdown at 1-based column / row k:
 *          LU:       max |pivot candidate| < eps       (scirs2-linalg/src/lapack.rs:106-120)
 *          Cholesky: a_ii - sum_{k<i} l_ik^2 <= 0       (scirs2-linalg/src/lapack.rs:661-670)
 *    k<0   SCIRS_ERR_* (message in scirs_gpu_last_error())
 *
 * Host-memory entry points copy their inputs to HBM, compute, and copy the
 * results back; they keep no pointer after returning and serialise on one
 * internal mutex (thread safe).  The *_dev entry points take HBM pointers and
 * a hipStream_t for data that is already resident (NULL = the default stream,
 * as everywhere in HIP: the work is ordered after what the caller queued on
 * it); they return after the stream has drained.
 */
#ifndef SCIRS_GPU_H
#define SCIRS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCIRS_OK 0
#define SCIRS_ERR_ARG (-1)
#define SCIRS_ERR_HIP (-2)
#define SCIRS_ERR_NOMEM (-3)
#define SCIRS_ERR_TIMEOUT (-4)
#define SCIRS_ERR_NOT_CONVERGED (-5)
#define SCIRS_ERR_UNSUPPORTED (-6)

/* solve paths reported in scirs_solve_info.path */
#define SCIRS_PATH_FP64_LU 0        /* plain FP64 LU + substitution            */
#define SCIRS_PATH_FP64_CHOL 1      /* plain FP64 Cholesky + substitution      */
#define SCIRS_PATH_MIXED_IR 2       /* low-precision LU + FP64 refinement       */
#define SCIRS_PATH_MIXED_GMRES 3    /* low-precision LU + FP64 GMRES-IR         */
#define SCIRS_PATH_FP64_FALLBACK 4  /* mixed path gave up; FP64 LU was used     */
#define SCIRS_PATH_FP32_LU 5
#define SCIRS_PATH_FP32_CHOL 6

/* precision modes of the mixed-precision solve (trailing-update arithmetic) */
#define SCIRS_MODE_AUTO 0           /* SCIRS_GPU_MIXED_MODE, else BF16X6        */
#define SCIRS_MODE_BF16 1           /* BF16 x BF16 -> FP32 matrix cores         */
#define SCIRS_MODE_FP16 2           /* FP16 with power-of-two scaling -> FP32   */
#define SCIRS_MODE_BF16X3 3         /* split BF16 (hi+lo), three products       */
#define SCIRS_MODE_FP32 4           /* FP32 arithmetic                          */
#define SCIRS_MODE_BF16X6 5         /* split BF16 (hi+mid+lo), six products:
                                       FP32-accurate on BF16 matrix cores       */
#define SCIRS_MODE_FP16X3 6         /* FP16 with power-of-two scaling, split
                                       (hi+lo), three products on FP16 cores    */

typedef struct scirs_solve_info {
    int32_t status;          /* same as the return value                          */
    int32_t path;            /* SCIRS_PATH_*                                      */
    int32_t ir_iters;        /* refinement steps (mixed path)                     */
    int32_t gmres_iters;     /* total GMRES inner iterations (mixed path)         */
    double pivot_value;      /* |pivot| or a_ii - sum at the breakdown (status>0) */
    double backward_error;   /* max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf); <0 = not computed */
    double time_factor_ms;   /* device time of the factorisation                  */
    double time_solve_ms;    /* device time of substitution / refinement          */
    double time_total_ms;    /* wall time of the call                             */
} scirs_solve_info;

/* ---- library ------------------------------------------------------------ */
/* Binds the library to a HIP device (default 0); called implicitly by every
 * entry point on first use. */
int scirs_gpu_init(int device);
/* Message of the last error on any thread ("" when none). */
const char *scirs_gpu_last_error(void);

/* ---- LU ------------------------------------------------------------------ */
/* Replaces scirs2-linalg/src/lapack.rs:72-147 `lu_factor` (called by
 * decomposition.rs:137 `lu`).  In place: a (m x n) receives the packed factors
 * (unit-lower L below the diagonal, U on and above); perm (m entries) receives
 * the reference's `piv` vector: row i of P*A is row perm[i] of A. */
int scirs_gpu_dgetrf(int m, int n, double *a, int64_t lda, int64_t *perm, scirs_solve_info *info);
int scirs_gpu_sgetrf(int m, int n, float *a, int64_t lda, int64_t *perm, scirs_solve_info *info);

/* Replaces the body of scirs2-linalg/src/solve.rs:54-111 `solve` (n > 4) and
 * solve.rs:366-424 `solve_multiple`: X = A^-1 B for B n x nrhs. */
int scirs_gpu_dgesv(int n, int nrhs, const double *a, int64_t lda, const double *b, int64_t ldb,
                    double *x, int64_t ldx, scirs_solve_info *info);
int scirs_gpu_sgesv(int n, int nrhs, const float *a, int64_t lda, const float *b, int64_t ldb,
                    float *x, int64_t ldx, scirs_solve_info *info);

/* ---- Cholesky -------------------------------------------------------------- */
/* Replaces scirs2-linalg/src/lapack.rs:632-679 `cholesky` (called by
 * decomposition.rs:44 `cholesky`).  Reads only the lower triangle of a (n x n)
 * and overwrites a with L (zeros above the diagonal). */
int scirs_gpu_dpotrf(int n, double *a, int64_t lda, scirs_solve_info *info);
int scirs_gpu_spotrf(int n, float *a, int64_t lda, scirs_solve_info *info);

/* The Cholesky-solve composition `cholesky` -> solve_triangular(L) ->
 * solve_triangular(L^T) (decomposition.rs:44, solve.rs:139): X = A^-1 B. */
int scirs_gpu_dposv(int n, int nrhs, const double *a, int64_t lda, const double *b, int64_t ldb,
                    double *x, int64_t ldx, scirs_solve_info *info);
/* The same composition for F = f32 (decomposition.rs:44 and solve.rs:139-202
 * instantiated at f32): factor and both substitutions compute in f32. */
int scirs_gpu_sposv(int n, int nrhs, const float *a, int64_t lda, const float *b, int64_t ldb,
                    float *x, int64_t ldx, scirs_solve_info *info);

/* ---- triangular ------------------------------------------------------------ */
/* Replaces scirs2-linalg/src/solve.rs:139-202 `solve_triangular` for nrhs
 * right-hand sides.  Returns i+1 when a non-unit diagonal has |a_ii| < eps. */
int scirs_gpu_dtrsv(int n, int nrhs, const double *a, int64_t lda, const double *b, int64_t ldb,
                    double *x, int64_t ldx, int lower, int unit_diagonal, scirs_solve_info *info);
/* solve.rs:139-202 `solve_triangular` for F = f32: substitution in f32, and the
 * singular test |a_ii| < f32::EPSILON (solve.rs:166, 186). */
int scirs_gpu_strsv(int n, int nrhs, const float *a, int64_t lda, const float *b, int64_t ldb,
                    float *x, int64_t ldx, int lower, int unit_diagonal, scirs_solve_info *info);

/* ---- QR and least squares --------------------------------------------------- */
/* Replaces scirs2-linalg/src/lapack.rs:171-253 `qr_factor` (behind
 * decomposition.rs:208 `qr`) for f64: a is rows x cols with rows >= cols >= 1;
 * q (rows x rows) and r (rows x cols) receive the reference's Householder
 * factors (unit reflectors, alpha = -sign(x0) ||x||, columns with ||x|| <= eps
 * skipped; r's sub-diagonal holds what the reflections leave, as in the
 * reference). */
int scirs_gpu_dqr(int rows, int cols, const double *a, int64_t lda, double *q, int64_t ldq, double *r,
                  int64_t ldr, scirs_solve_info *info);

/* Replaces the rows >= cols branch of scirs2-linalg/src/solve.rs:234-290
 * `lstsq` for f64: x (cols) minimises ||b - A x||_2 through R x = (Q^T b)[..cols],
 * *residual = ||b - A x||_2^2.  Returns k+1 when the back substitution meets
 * |r_kk| < eps (solve.rs:184-191), bottom row first. */
int scirs_gpu_dlstsq(int rows, int cols, const double *a, int64_t lda, const double *b, double *x,
                     double *residual, scirs_solve_info *info);

/* ---- Krylov ------------------------------------------------------------------- */
/* Replaces scirs2-linalg/src/solvers/iterative.rs:317-450 `gmres` for f64:
 * restarted GMRES with the reference's classical Gram-Schmidt, breakdown test
 * (h_{j+1,j} < eps ends the cycle without counting an iteration), cycle count
 * (max_iter / restart).max(1) and Givens least-squares solve (:604-664).
 * x0 may be NULL (zeros); restart <= 0 means min(n, 30).  Outputs the solution,
 * the iteration count, ||b - A x||_2 and whether it fell below tol.
 * Returns 0 (converged or not), or k > 0 when the least-squares solve meets
 * |r_kk| < eps (the reference's SingularMatrixError). */
int scirs_gpu_dgmres(int n, const double *a, int64_t lda, const double *b, const double *x0, int max_iter,
                     double tol, int restart, double *x, int *iterations, double *residual_norm, int *converged,
                     scirs_solve_info *info);
/* The same on device-resident A (lda), b and x, on the caller's stream (NULL =
 * the default stream): x holds x0 when use_x0 != 0 (else it is zeroed first)
 * and the solution on return.  The host reads a few scalars per step (the
 * reference's control flow), so the call returns after the work is done. */
int scirs_gpu_dgmres_dev(int n, const double *dA, int64_t lda, const double *dB, double *dX, int use_x0,
                         int max_iter, double tol, int restart, int *iterations, double *residual_norm,
                         int *converged, void *stream, scirs_solve_info *info);

/* Replaces scirs2-linalg/src/preconditioners.rs:934-1010 `preconditioned_gmres`
 * for f64 with a real solver (the reference's body never solves its
 * least-squares problem and ends with x += 0.1 tol v_0): the gmres iteration
 * above, right-preconditioned — the Arnoldi process runs on A M^-1 and each
 * cycle ends with x += M^-1 (V y) — so residual norm and stopping test are
 * those of A x = b.  precond: SCIRS_PRECOND_NONE, _DIAGONAL (the Jacobi
 * preconditioner of preconditioners.rs:210-229: 1/a_ii, or 1 where
 * |a_ii| < eps), _LU (the library factors A: FP32 tournament-pivoted panel,
 * FP32-accurate split-BF16 update, applied in FP64), _LU_GIVEN (lu: the
 * caller's n x n packed FP64 factors, unit lower L below the diagonal and U on
 * and above, with perm[i] the row of A at position i: P A = L U).  Returns as
 * scirs_gpu_dgmres; with _LU, k > 0 also when the FP32 factorisation meets a
 * zero pivot at column k. */
#define SCIRS_PRECOND_NONE 0
#define SCIRS_PRECOND_DIAGONAL 1
#define SCIRS_PRECOND_LU 2
#define SCIRS_PRECOND_LU_GIVEN 3
int scirs_gpu_dgmres_precond(int n, const double *a, int64_t lda, const double *b, const double *x0, int max_iter,
                             double tol, int restart, int precond, const double *lu, const int64_t *perm, double *x,
                             int *iterations, double *residual_norm, int *converged, scirs_solve_info *info);
/* The same on device-resident data; dLU / dPerm (device) for _LU_GIVEN. */
int scirs_gpu_dgmres_precond_dev(int n, const double *dA, int64_t lda, const double *dB, double *dX, int use_x0,
                                 int max_iter, double tol, int restart, int precond, const double *dLU,
                                 const int64_t *dPerm, int *iterations, double *residual_norm, int *converged,
                                 void *stream, scirs_solve_info *info);

/* ---- batched array operations ------------------------------------------------ */
/* Replace scirs2-linalg/src/batch/mod.rs for f64 / f32, bit-identical to its
 * loops (each element summed from +0 in the reference's order, no fused
 * multiply-add).  Batches are contiguous C-order arrays batch x m x k.
 *   batch_matmul (mod.rs:62-107):  c[p] = a[p] b,  b is k x n, c batch x m x n
 *   batch_matvec (mod.rs:139-172): y[p] = a[p] x,  a batch x m x n, y batch x m
 *   batch_add    (mod.rs:220-291): out = a + v along axis 0 (v[i], len m) or
 *                                  axis 1 (v[j], len n); other axes -> SCIRS_ERR_ARG
 *   batch_sum    (mod.rs:326-336): out (m x n) = sum over the batch, in batch order
 * Empty shapes return 0 and write nothing (k = 0 gives zeros). */
int scirs_gpu_dbatch_matmul(int batch, int m, int k, int n, const double *a, const double *b, double *c,
                            scirs_solve_info *info);
int scirs_gpu_sbatch_matmul(int batch, int m, int k, int n, const float *a, const float *b, float *c,
                            scirs_solve_info *info);
int scirs_gpu_dbatch_matvec(int batch, int m, int n, const double *a, const double *x, double *y,
                            scirs_solve_info *info);
int scirs_gpu_sbatch_matvec(int batch, int m, int n, const float *a, const float *x, float *y,
                            scirs_solve_info *info);
int scirs_gpu_dbatch_add(int batch, int m, int n, const double *a, const double *v, int axis, double *out,
                         scirs_solve_info *info);
int scirs_gpu_sbatch_add(int batch, int m, int n, const float *a, const float *v, int axis, float *out,
                         scirs_solve_info *info);
int scirs_gpu_dbatch_sum(int batch, int m, int n, const double *a, double *out, scirs_solve_info *info);
int scirs_gpu_sbatch_sum(int batch, int m, int n, const float *a, float *out, scirs_solve_info *info);
/* batch_matmul on device-resident arrays, on the caller's stream (NULL = the
 * default stream); no synchronisation. */
int scirs_gpu_dbatch_matmul_dev(int batch, int m, int k, int n, const double *dA, const double *dB, double *dC,
                                void *stream);

/* ---- mixed precision ---------------------------------------------------------- */
/* The factor-once counterpart of scirs2-linalg/src/mixed_precision/adaptive.rs:282-378
 * `iterative_refinement_solve::<f64,f64,f64,f64,f32>`: FP32 panel, low-precision
 * trailing update (mode), FP64 residual on the GPU, refinement until the
 * normwise backward error is <= tol (tol <= 0: n * 2^-53), GMRES-IR
 * (restarted, preconditioned by the low-precision factors) when refinement
 * stalls, and FP64 LU when both fail.  max_iter / max_gmres <= 0 take the
 * defaults (30 refinement steps; 300 GMRES iterations per right-hand side). */
int scirs_gpu_dgesv_ir(int n, int nrhs, const double *a, int64_t lda, const double *b, int64_t ldb,
                       double *x, int64_t ldx, int mode, int max_iter, int max_gmres, double tol,
                       scirs_solve_info *info);

/* ---- device-resident variants ---------------------------------------------- */
/* dA is not modified; dLU (n x n, ld ldlu) receives the factors and dPerm
 * (n) the permutation; dX receives the solution.  dLU / dPerm may be NULL
 * (library scratch is used). */
int scirs_gpu_dgesv_dev(int n, int nrhs, const double *dA, int64_t lda, const double *dB, int64_t ldb,
                        double *dX, int64_t ldx, double *dLU, int64_t ldlu, int64_t *dPerm,
                        void *stream, scirs_solve_info *info);
int scirs_gpu_dposv_dev(int n, int nrhs, const double *dA, int64_t lda, const double *dB, int64_t ldb,
                        double *dX, int64_t ldx, double *dL, int64_t ldl, void *stream,
                        scirs_solve_info *info);
int scirs_gpu_dgesv_ir_dev(int n, int nrhs, const double *dA, int64_t lda, const double *dB, int64_t ldb,
                           double *dX, int64_t ldx, int mode, int max_iter, int max_gmres, double tol,
                           void *stream, scirs_solve_info *info);

/* ---- batched uneven solves ------------------------------------------------------ */
/* No reference counterpart: scirs2-linalg/src/batch/mod.rs has no batched solve.
 * Each system i follows the reference's per-system semantics — `solve`
 * (solve.rs:54-111) when spd[i] == 0, the cholesky + substitution composition
 * (decomposition.rs:44, solve.rs:139) when spd[i] != 0 — computed as an FP32
 * factorisation plus FP64 refinement to a normwise backward error <= tol_scale *
 * n_i * 2^-53 (FP64 factorisation when FP32 breaks down or stalls), all systems in
 * one persistent, hipGraph-replayed launch.
 * Host arrays: n, nrhs (1 <= n <= 256, 1 <= nrhs <= 16), spd (may be NULL), offA,
 * offB (element offsets of A_i, row-major n_i x n_i, and of B_i / X_i, row-major
 * n_i x nrhs_i), status (out: 0, or the breakdown column / row), bwd (out, may be
 * NULL).  Device arrays: dA, dB, dX (packed). */
typedef struct scirs_batch_info {
    int32_t failed;              /* systems with status != 0                         */
    int32_t max_iters;           /* most refinement steps taken by any system        */
    double max_backward_error;   /* over the systems with status 0                   */
    double time_ms;              /* device time of the batched launch                */
    double total_ms;             /* wall time of the call                            */
    int32_t wave_max;            /* systems with n <= wave_max ran one per wave      */
    int32_t graph_replayed;      /* 1 when the launch replayed a cached hipGraph     */
} scirs_batch_info;
int scirs_gpu_dgesv_batched_dev(int batch, const int *n, const int *nrhs, const unsigned char *spd,
                                const double *dA, const int64_t *offA, const double *dB, const int64_t *offB,
                                double *dX, int *status, double *bwd, int max_iter, double tol_scale, void *stream,
                                scirs_batch_info *info);
/* The same batched solve on host memory (the entry a Rust `solve_batched` binds
 * without HIP calls of its own): A[i] (n_i x n_i, row stride lda[i]), B[i] and
 * X[i] (n_i x nrhs_i, row strides ldb[i], ldx[i]) are host pointers.  The
 * systems travel in one pinned copy each way; status, bwd and info as above
 * (info->total_ms includes both copies).  X[i] is written for every system;
 * it holds no solution where status[i] != 0. */
int scirs_gpu_dgesv_batched(int batch, const int *n, const int *nrhs, const unsigned char *spd,
                            const double *const *A, const int64_t *lda, const double *const *B, const int64_t *ldb,
                            double *const *X, const int64_t *ldx, int *status, double *bwd, int max_iter,
                            double tol_scale, scirs_batch_info *info);

/* ---- diagnostics -------------------------------------------------------------- */
/* Kernel-class timing with HIP events on the launching stream (no reference
 * counterpart; the reference's profiling layer, scirs2-core/src/profiling.rs,
 * is not wired into linalg).  Classes, in array order: 0 panel, 1 row
 * interchanges, 2 TRSM, 3 trailing update, 4 substitution, 5 residual,
 * 6 conversion, 7 other.  scirs_gpu_prof_end fills SCIRS_PROF_NCAT records. */
#define SCIRS_PROF_NCAT 8
typedef struct scirs_prof_stats {
    int64_t launches;    /* timed scopes (one scope may cover several launches) */
    double total_ms;     /* summed event time                                    */
    double flops;        /* algorithmic floating-point operations                */
    double bytes;        /* algorithmic HBM bytes                                */
} scirs_prof_stats;
int scirs_gpu_prof_begin(void);
int scirs_gpu_prof_end(scirs_prof_stats *out);

/* Normwise backward error of X for A X = B, all in HBM (FP64 residual on the
 * GPU).  No reference counterpart: the reference's tests compute residuals
 * with ndarray `dot` (tests/scipy_compat_validation.rs:455-476). */
int scirs_gpu_backward_error_dev(int n, int nrhs, const double *dA, int64_t lda, const double *dX, int64_t ldx,
                                 const double *dB, int64_t ldb, void *stream, double *out);

/* The same backward error with no allocation: its scratch (the residual, the
 * row sums of |A| and the norms: n * nrhs + n + 200 doubles) is a fixed 8 MiB
 * array in the library's code object instead of memory the library allocates,
 * so the call makes no hipMalloc / hipFree or other allocating call.  No
 * counterpart in the reference; it is for callers that cannot allocate (code
 * being captured into a HIP graph, tight serving loops, harnesses that forbid
 * allocation).  Limits: n * nrhs + n + 200 <= 1048576 (else SCIRS_ERR_ARG);
 * scirs_gpu_init must have run (it resolves the array's device address;
 * otherwise SCIRS_ERR_UNSUPPORTED).  Every call on the device shares the one
 * array: calls that overlap on the device race (the library's lock serialises
 * the calls of one process, each of which waits for its stream). */
int scirs_gpu_backward_error_dev_noalloc(int n, int nrhs, const double *dA, int64_t lda, const double *dX,
                                         int64_t ldx, const double *dB, int64_t ldb, void *stream, double *out);

/* Tournament-pivoted FP32 LU (CALU) of the n x n matrix in HBM, in place (unit L
 * below the diagonal, U on and above; lda a multiple of 4), with the trailing
 * update in the arithmetic of `mode` (SCIRS_MODE_FP32, _BF16, _BF16X3, _BF16X6;
 * AUTO = BF16X6), and the row permutation in perm (int64, row i of P A is row
 * perm[i] of A).  No counterpart in the reference: this is the factorisation
 * scirs_gpu_dgesv_ir refines (its pivots are chosen by a reduction tree over
 * blocks of rows, not the column arg-max of lapack.rs:95-104), exported so its
 * factors can be checked on their own.  Returns 0, k > 0 when |u_kk| < eps
 * (FP32), or SCIRS_ERR_*. */
int scirs_gpu_sgetrf_tp_dev(int n, float *dA, int64_t lda, int64_t *dperm, int mode, void *stream);

/* Host arithmetic of the exchange tags (no device work, no reference
 * counterpart; exported so the wrap-around rule can be tested on the host).
 * A launch that uses `span` consecutive tags gets base .. base + span - 1 from
 * the running count `next`; *next_out is the count after it.  Returns 1 when the
 * count had to restart at 1 (the library then clears every tagged buffer
 * first), 0 otherwise.  Tag 0 is never handed out. */
int scirs_gpu_seq_plan(unsigned next, unsigned span, unsigned *base, unsigned *next_out);

#ifdef __cplusplus
}
#endif

#endif /* SCIRS_GPU_H */
