/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the CPU algorithms that scirs2-linalg uses on the
 * dense-solve path.  Only the test suite, the smoke check and the CPU
 * baseline leg of the benchmark may load this file's shared library; the
 * product path (libscirs_gpu.so) never links or calls it.
 *
 * Each function follows the reference's operation order so that, compiled
 * with -O2 -ffp-contract=off (Rust never fuses a*b+c), its results are the
 * ones the Rust code produces bit for bit:
 *
 *   oracle_lu_factor_*        scirs2-linalg/src/lapack.rs:72-147
 *   oracle_cholesky_*         scirs2-linalg/src/lapack.rs:632-679
 *   oracle_solve_triangular_* scirs2-linalg/src/solve.rs:139-202
 *   oracle_solve_*            scirs2-linalg/src/solve.rs:54-111
 *                             (n <= 4 branch: basic.rs:117-186 inv, basic.rs:29-93 det)
 *   oracle_solve_multiple_*   scirs2-linalg/src/solve.rs:366-424
 *   oracle_qr_factor_d        scirs2-linalg/src/lapack.rs:171-253
 *   oracle_lstsq_d            scirs2-linalg/src/solve.rs:234-290 (rows >= cols)
 *   oracle_batch_*            scirs2-linalg/src/batch/mod.rs:62-336
 *   oracle_gmres_d            scirs2-linalg/src/solvers/iterative.rs:317-450, 604-664
 *
 * All matrices are dense row-major with an explicit leading dimension, the
 * layout of an ndarray in standard order.
 *
 * Return codes: 0 = success; k > 0 = singular / not positive definite, with
 * k the 1-based column (LU) or row (Cholesky) at which the reference raises
 * its error; -1 = empty input; -2 = dimension error; -3 = memory.
 *
 * Pinning (which check fixes which function) is listed in the header of
 * tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define ORACLE_SINGULAR_COND_NONE (-1.0)

/* Instantiate every routine for one scalar type.  T = element type,
 * S = name suffix, EPS = T's machine epsilon (num_traits Float::epsilon). */
#define ORACLE_DEFINE(T, S, EPS, FABS, SQRT)                                        \
                                                                                    \
/* lapack.rs:72-147.  lu (m_rows x n_cols, row-major, ld = n_cols) receives the   \
 * packed factors, piv the permutation vector (row i of P*A is row piv[i] of A),   \
 * sign the permutation sign and cond_est the reference's 1/max estimate when the  \
 * pivot is tiny but non-zero (else -1). */                                        \
int oracle_lu_factor_##S(int n, int m, const T *a, long lda, T *lu, int64_t *piv,  \
                         T *sign, double *cond_est)                                \
{                                                                                   \
    if (n == 0 || m == 0) return -1; /* lapack.rs:79-83 */                          \
    for (int i = 0; i < n; ++i)                                                     \
        memcpy(lu + (size_t)i * m, a + (size_t)i * lda, sizeof(T) * (size_t)m);     \
    for (int i = 0; i < n; ++i) piv[i] = i;                                         \
    T sg = (T)1;                                                                    \
    if (cond_est) *cond_est = ORACLE_SINGULAR_COND_NONE;                            \
    int kmax = n < m ? n : m;                                                       \
    for (int k = 0; k < kmax; ++k) {                                                \
        /* pivot search, strict '>' so the first maximum wins (lapack.rs:95-104) */ \
        int p = k;                                                                  \
        T max_val = FABS(lu[(size_t)k * m + k]);                                    \
        for (int i = k + 1; i < n; ++i) {                                           \
            T av = FABS(lu[(size_t)i * m + k]);                                     \
            if (av > max_val) { max_val = av; p = i; }                              \
        }                                                                           \
        /* absolute singularity test (lapack.rs:106-120) */                         \
        if (max_val < EPS) {                                                        \
            if (cond_est && max_val > (T)0) *cond_est = (double)((T)1 / max_val);   \
            if (sign) *sign = sg;                                                   \
            return k + 1;                                                           \
        }                                                                           \
        /* full-row swap (lapack.rs:123-134) */                                     \
        if (p != k) {                                                               \
            for (int j = 0; j < m; ++j) {                                           \
                T t = lu[(size_t)k * m + j];                                        \
                lu[(size_t)k * m + j] = lu[(size_t)p * m + j];                      \
                lu[(size_t)p * m + j] = t;                                          \
            }                                                                       \
            int64_t tp = piv[k]; piv[k] = piv[p]; piv[p] = tp;                      \
            sg = -sg;                                                               \
        }                                                                           \
        /* multipliers by true division, then the row update (lapack.rs:137-143) */ \
        for (int i = k + 1; i < n; ++i) {                                           \
            T *ri = lu + (size_t)i * m;                                             \
            const T *rk = lu + (size_t)k * m;                                       \
            ri[k] = ri[k] / rk[k];                                                  \
            T l = ri[k];                                                            \
            for (int j = k + 1; j < m; ++j) ri[j] = ri[j] - l * rk[j];              \
        }                                                                           \
    }                                                                               \
    if (sign) *sign = sg;                                                           \
    return 0;                                                                       \
}                                                                                   \
                                                                                    \
/* lapack.rs:632-679, Cholesky-Banachiewicz.  Reads only the lower triangle of a;  \
 * l (n x n, ld = n) gets the factor with zeros above the diagonal.  Returns i+1   \
 * when a_ii - sum <= 0 at row i. */                                               \
int oracle_cholesky_##S(int n, const T *a, long lda, T *l)                          \
{                                                                                   \
    if (n == 0) return -1;                                                          \
    memset(l, 0, sizeof(T) * (size_t)n * (size_t)n);                                \
    for (int i = 0; i < n; ++i) {                                                   \
        for (int j = 0; j <= i; ++j) {                                              \
            T sum = (T)0;                                                           \
            for (int k = 0; k < j; ++k)                                             \
                sum += l[(size_t)i * n + k] * l[(size_t)j * n + k];                 \
            if (i == j) {                                                           \
                T val = a[(size_t)i * lda + i] - sum;                               \
                if (val <= (T)0) return i + 1;                                      \
                l[(size_t)i * n + j] = SQRT(val);                                   \
            } else {                                                                \
                l[(size_t)i * n + j] = (a[(size_t)i * lda + j] - sum) /             \
                                       l[(size_t)j * n + j];                        \
            }                                                                       \
        }                                                                           \
    }                                                                               \
    return 0;                                                                       \
}                                                                                   \
                                                                                    \
/* solve.rs:139-202 (the numeric part; the finiteness checks live in the host     \
 * layer).  Returns i+1 when |a_ii| < eps on a non-unit diagonal. */               \
int oracle_solve_triangular_##S(int n, const T *a, long lda, const T *b, T *x,      \
                                int lower, int unit_diagonal)                       \
{                                                                                   \
    if (n == 0) return -1;                                                          \
    if (lower) {                                                                    \
        for (int i = 0; i < n; ++i) {                                               \
            T sum = b[i];                                                           \
            for (int j = 0; j < i; ++j) sum -= a[(size_t)i * lda + j] * x[j];       \
            if (unit_diagonal) {                                                    \
                x[i] = sum;                                                         \
            } else {                                                                \
                T d = a[(size_t)i * lda + i];                                       \
                if (FABS(d) < EPS) return i + 1;                                    \
                x[i] = sum / d;                                                     \
            }                                                                       \
        }                                                                           \
    } else {                                                                        \
        for (int i = n - 1; i >= 0; --i) {                                          \
            T sum = b[i];                                                           \
            for (int j = i + 1; j < n; ++j) sum -= a[(size_t)i * lda + j] * x[j];   \
            if (unit_diagonal) {                                                    \
                x[i] = sum;                                                         \
            } else {                                                                \
                T d = a[(size_t)i * lda + i];                                       \
                if (FABS(d) < EPS) return i + 1;                                    \
                x[i] = sum / d;                                                     \
            }                                                                       \
        }                                                                           \
    }                                                                               \
    return 0;                                                                       \
}                                                                                   \
                                                                                    \
/* LU-based solve of one column, the body of solve.rs:84-110 and of the per-column \
 * loop solve.rs:399-421.  lu/piv come from oracle_lu_factor.  P*b is exactly     \
 * b[piv[i]] (a sum of one b entry and zeros, solve.rs:96-102). */                 \
static int oracle_lu_apply_##S(int n, const T *lu, const int64_t *piv,             \
                               const T *b, long bstride, T *x, long xstride,        \
                               T *work)                                             \
{                                                                                   \
    T *pb = work, *y = work + n, *xx = work + 2 * n;                                \
    for (int i = 0; i < n; ++i) pb[i] = b[(size_t)piv[i] * bstride];                \
    /* L is the strict lower part with a unit diagonal (decomposition.rs:165-173) */\
    oracle_solve_triangular_##S(n, lu, n, pb, y, 1, 1);                             \
    int rc = oracle_solve_triangular_##S(n, lu, n, y, xx, 0, 0);                    \
    if (rc) return rc;                                                              \
    for (int i = 0; i < n; ++i) x[(size_t)i * xstride] = xx[i];                     \
    return 0;                                                                       \
}                                                                                   \
                                                                                    \
/* solve.rs:366-424.  b is n x nrhs (ld ldb), x is n x nrhs (ld ldx). */            \
int oracle_solve_multiple_##S(int n, int nrhs, const T *a, long lda, const T *b,    \
                              long ldb, T *x, long ldx)                             \
{                                                                                   \
    if (n == 0 || nrhs == 0) return -1;                                             \
    T *lu = (T *)malloc(sizeof(T) * (size_t)n * (size_t)n);                         \
    int64_t *piv = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);                  \
    T *work = (T *)malloc(sizeof(T) * 3 * (size_t)n);                               \
    if (!lu || !piv || !work) { free(lu); free(piv); free(work); return -3; }       \
    int rc = oracle_lu_factor_##S(n, n, a, lda, lu, piv, NULL, NULL);               \
    for (int j = 0; rc == 0 && j < nrhs; ++j)                                       \
        rc = oracle_lu_apply_##S(n, lu, piv, b + j, ldb, x + j, ldx, work);         \
    free(lu); free(piv); free(work);                                                \
    return rc;                                                                      \
}                                                                                   \
                                                                                    \
/* basic.rs:117-186 for n <= 4: the 2x2 closed form with its |det| < eps test     \
 * (det from basic.rs:47), otherwise solve_multiple(A, I). */                      \
static int oracle_inv_small_##S(int n, const T *a, long lda, T *ainv)               \
{                                                                                   \
    if (n == 2) {                                                                   \
        T a00 = a[0], a01 = a[1], a10 = a[lda], a11 = a[lda + 1];                   \
        T det_val = a00 * a11 - a01 * a10;                                          \
        if (FABS(det_val) < EPS) return 1;                                          \
        T inv_det = (T)1 / det_val;                                                 \
        ainv[0] = a11 * inv_det;                                                    \
        ainv[1] = -a01 * inv_det;                                                   \
        ainv[2] = -a10 * inv_det;                                                   \
        ainv[3] = a00 * inv_det;                                                    \
        return 0;                                                                   \
    }                                                                               \
    T eye[16];                                                                      \
    memset(eye, 0, sizeof(eye));                                                    \
    for (int i = 0; i < n; ++i) eye[i * n + i] = (T)1;                              \
    return oracle_solve_multiple_##S(n, n, a, lda, eye, n, ainv, n);                \
}                                                                                   \
                                                                                    \
/* solve.rs:54-111.  For n <= 4 the reference forms inv(A) and multiplies         \
 * (solve.rs:65-76); above that it factors and substitutes. */                     \
int oracle_solve_##S(int n, const T *a, long lda, const T *b, T *x)                \
{                                                                                   \
    if (n == 0) return -1;                                                          \
    if (n <= 4) {                                                                   \
        T ainv[16];                                                                 \
        int rc = oracle_inv_small_##S(n, a, lda, ainv);                             \
        if (rc) return rc;                                                          \
        for (int i = 0; i < n; ++i) {                                               \
            T acc = (T)0;                                                           \
            for (int j = 0; j < n; ++j) acc += ainv[i * n + j] * b[j];              \
            x[i] = acc;                                                             \
        }                                                                           \
        return 0;                                                                   \
    }                                                                               \
    return oracle_solve_multiple_##S(n, 1, a, lda, b, 1, x, 1);                     \
}

ORACLE_DEFINE(double, d, DBL_EPSILON, fabs, sqrt)
ORACLE_DEFINE(float, s, FLT_EPSILON, fabsf, sqrtf)

/* lapack.rs:171-253 (qr_factor), f64.  a is rows x cols (rows >= cols), q is
 * rows x rows, r is rows x cols.  Householder reflectors with v normalised to
 * unit length (H = I - 2 v v^T), alpha = -sign(x0)*||x|| (x0 >= 0 -> -||x||),
 * a column skipped when ||x|| <= eps or ||v|| <= eps.  Q is accumulated as
 * Q^T (reflectors applied to the identity from the left) and transposed at
 * the end; sub-diagonal entries of R keep whatever the reflections leave.
 * Returns -1 for an empty matrix, -2 for rows < cols. */
int oracle_qr_factor_d(int rows, int cols, const double *a, long lda, double *q, double *r)
{
    if (rows == 0 || cols == 0) return -1;
    if (rows < cols) return -2;
    const int n = rows, m = cols;
    double *qt = (double *)malloc(sizeof(double) * (size_t)n * (size_t)n);
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    double *col = (double *)malloc(sizeof(double) * (size_t)n);
    if (!qt || !v || !col) { free(qt); free(v); free(col); return -3; }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) r[(size_t)i * m + j] = a[(size_t)i * lda + j];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) qt[(size_t)i * n + j] = i == j ? 1.0 : 0.0;
    for (int k = 0; k < (m < n ? m : n); ++k) {
        const int len = n - k;
        double xs = 0.0;
        for (int i = 0; i < len; ++i) {
            v[i] = r[(size_t)(k + i) * m + k];
            xs += v[i] * v[i];
        }
        const double x_norm = sqrt(xs);
        if (!(x_norm > DBL_EPSILON)) continue;
        const double alpha = v[0] >= 0.0 ? -x_norm : x_norm;
        v[0] -= alpha;
        double vs = 0.0;
        for (int i = 0; i < len; ++i) vs += v[i] * v[i];
        const double v_norm = sqrt(vs);
        if (!(v_norm > DBL_EPSILON)) continue;
        for (int i = 0; i < len; ++i) v[i] /= v_norm;
        for (int j = k; j < m; ++j) {
            double dot = 0.0;
            for (int i = 0; i < len; ++i) dot = dot + v[i] * r[(size_t)(k + i) * m + j];
            for (int i = 0; i < len; ++i) r[(size_t)(k + i) * m + j] -= 2.0 * v[i] * dot;
        }
        for (int j = 0; j < n; ++j) {
            double dot = 0.0;
            for (int i = 0; i < len; ++i) col[i] = qt[(size_t)(k + i) * n + j];
            for (int i = 0; i < len; ++i) dot = dot + v[i] * col[i];
            for (int i = 0; i < len; ++i) qt[(size_t)(k + i) * n + j] -= 2.0 * v[i] * dot;
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) q[(size_t)i * n + j] = qt[(size_t)j * n + i];
    free(qt); free(v); free(col);
    return 0;
}

/* solve.rs:234-290, the rows >= cols branch of lstsq, f64: (Q, R) = qr(A),
 * Q^T b with Q^T formed explicitly and summed in order, back substitution on
 * the leading cols x cols block of R (solve.rs:139-202, its |r_ii| < eps test
 * returning the 1-based row), then the residual sum of squares ||b - A x||^2
 * accumulated row by row.  Returns what qr_factor or the substitution returns. */
int oracle_lstsq_d(int rows, int cols, const double *a, long lda, const double *b, double *x,
                   double *residual)
{
    if (rows == 0 || cols == 0) return -1;
    if (rows < cols) return -2;
    const int n = rows, m = cols;
    double *q = (double *)malloc(sizeof(double) * (size_t)n * (size_t)n);
    double *r = (double *)malloc(sizeof(double) * (size_t)n * (size_t)m);
    double *qtb = (double *)malloc(sizeof(double) * (size_t)n);
    if (!q || !r || !qtb) { free(q); free(r); free(qtb); return -3; }
    int rc = oracle_qr_factor_d(rows, cols, a, lda, q, r);
    if (rc == 0) {
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int j = 0; j < n; ++j) acc += q[(size_t)j * n + i] * b[j];
            qtb[i] = acc;
        }
        rc = oracle_solve_triangular_d(m, r, m, qtb, x, 0, 0);
    }
    if (rc == 0) {
        double res = 0.0;
        for (int i = 0; i < n; ++i) {
            double ax = 0.0;
            for (int j = 0; j < m; ++j) ax += a[(size_t)i * lda + j] * x[j];
            const double d = b[i] - ax;
            res += d * d;
        }
        *residual = res;
    }
    free(q); free(r); free(qtb);
    return rc;
}

/* batch/mod.rs:62-336.  batch a is batch x m x k (C order, contiguous).
 *   batch_matmul: c[p,i,j] = sum_k a[p,i,k] * b[k,j], summed from zero in k order
 *   batch_matvec: y[p,i]   = sum_j a[p,i,j] * x[j], summed from zero in j order
 *   batch_add:    a[p,i,j] + v[i] (axis 0) or + v[j] (axis 1)
 *   batch_sum:    sum over p of a[p,i,j], accumulated from zero in p order
 *                 (ndarray's sum_axis over a non-minimal-stride axis adds the
 *                 sub-views one after another) */
#define ORACLE_BATCH(T, S)                                                          \
void oracle_batch_matmul_##S(int batch, int m, int k, int n, const T *a,            \
                             const T *b, T *c)                                      \
{                                                                                   \
    for (int p = 0; p < batch; ++p)                                                 \
        for (int i = 0; i < m; ++i)                                                 \
            for (int j = 0; j < n; ++j) {                                           \
                T sum = (T)0;                                                       \
                const T *ar = a + ((size_t)p * m + i) * k;                          \
                for (int q = 0; q < k; ++q) sum += ar[q] * b[(size_t)q * n + j];    \
                c[((size_t)p * m + i) * n + j] = sum;                               \
            }                                                                       \
}                                                                                   \
void oracle_batch_matvec_##S(int batch, int m, int n, const T *a, const T *x, T *y)  \
{                                                                                   \
    for (int p = 0; p < batch; ++p)                                                 \
        for (int i = 0; i < m; ++i) {                                               \
            T sum = (T)0;                                                           \
            const T *ar = a + ((size_t)p * m + i) * n;                              \
            for (int j = 0; j < n; ++j) sum += ar[j] * x[j];                        \
            y[(size_t)p * m + i] = sum;                                             \
        }                                                                           \
}                                                                                   \
void oracle_batch_add_##S(int batch, int m, int n, const T *a, const T *v, int axis, \
                          T *out)                                                   \
{                                                                                   \
    for (int p = 0; p < batch; ++p)                                                 \
        for (int i = 0; i < m; ++i)                                                 \
            for (int j = 0; j < n; ++j) {                                           \
                const size_t o = ((size_t)p * m + i) * n + j;                       \
                out[o] = a[o] + (axis == 0 ? v[i] : v[j]);                          \
            }                                                                       \
}                                                                                   \
void oracle_batch_sum_##S(int batch, int m, int n, const T *a, T *out)              \
{                                                                                   \
    for (size_t o = 0; o < (size_t)m * n; ++o) out[o] = (T)0;                        \
    for (int p = 0; p < batch; ++p)                                                 \
        for (size_t o = 0; o < (size_t)m * n; ++o) out[o] = out[o] + a[(size_t)p * m * n + o]; \
}

ORACLE_BATCH(double, d)
ORACLE_BATCH(float, s)

/* solvers/iterative.rs:604-664 (solve_least_squares_gmres): Givens rotations
 * on the (m x nc) Hessenberg h (row-major, ld nc), then back substitution.
 * Returns 0, or i+1 when |r_ii| < eps (the reference's SingularMatrixError). */
static int oracle_gmres_lsq(int m, int nc, const double *h, double beta, double *y)
{
    double *hc = (double *)malloc(sizeof(double) * (size_t)m * (nc > 0 ? nc : 1));
    double *g = (double *)calloc((size_t)m, sizeof(double));
    if (!hc || !g) { free(hc); free(g); return -3; }
    memcpy(hc, h, sizeof(double) * (size_t)m * nc);
    g[0] = beta;
    for (int j = 0; j < nc; ++j)
        for (int i = j + 1; i < m; ++i) {
            if (fabs(hc[i * nc + j]) > DBL_EPSILON) {
                const double a = hc[j * nc + j], b = hc[i * nc + j];
                const double r = sqrt(a * a + b * b);
                const double c = a / r, s = b / r;
                for (int k = j; k < nc; ++k) {
                    const double hjk = hc[j * nc + k], hik = hc[i * nc + k];
                    hc[j * nc + k] = c * hjk + s * hik;
                    hc[i * nc + k] = -s * hjk + c * hik;
                }
                const double gj = g[j], gi = g[i];
                g[j] = c * gj + s * gi;
                g[i] = -s * gj + c * gi;
            }
        }
    int rc = 0;
    for (int i = nc - 1; i >= 0; --i) {
        double sum = g[i];
        for (int j = i + 1; j < nc; ++j) sum -= hc[i * nc + j] * y[j];
        if (fabs(hc[i * nc + i]) < DBL_EPSILON) { rc = i + 1; break; }
        y[i] = sum / hc[i * nc + i];
    }
    free(hc); free(g);
    return rc;
}

static double oracle_resid_norm(int n, const double *a, long lda, const double *b, const double *x, double *r)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        double ax = 0.0;
        for (int j = 0; j < n; ++j) ax += a[(size_t)i * lda + j] * x[j];
        r[i] = b[i] - ax;
    }
    for (int i = 0; i < n; ++i) s += r[i] * r[i];
    return sqrt(s);
}

/* M^-1 v of the right-preconditioned variant below: kind 0 identity, 1 the
 * Jacobi preconditioner of preconditioners.rs:210-229 (1/a_ii, or 1 where
 * |a_ii| < eps), 3 given LU factors (packed: unit lower L below the diagonal,
 * U on and above; P A = L U with perm[i] the original row at position i). */
static void oracle_prec_apply(int n, int kind, const double *dinv, const double *lu, const int64_t *perm,
                              const double *v, double *z)
{
    if (kind == 1) {
        for (int i = 0; i < n; ++i) z[i] = dinv[i] * v[i];
    } else if (kind == 3) {
        for (int i = 0; i < n; ++i) z[i] = v[perm[i]];
        for (int i = 0; i < n; ++i) {
            double s = z[i];
            for (int j = 0; j < i; ++j) s -= lu[(size_t)i * n + j] * z[j];
            z[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = z[i];
            for (int j = i + 1; j < n; ++j) s -= lu[(size_t)i * n + j] * z[j];
            z[i] = s / lu[(size_t)i * n + i];
        }
    } else {
        memcpy(z, v, sizeof(double) * (size_t)n);
    }
}

/* solvers/iterative.rs:317-450 (gmres), f64: restarted GMRES with classical
 * Gram-Schmidt against the unmodified w, a break (counting no iteration) when
 * h[j+1][j] < eps, (max_iterations / restart).max(1) cycles.  x0 may be NULL
 * (zeros); restart <= 0 means min(n, 30).  Returns 0, or k+1 > 0 when the
 * least-squares solve is singular at row k+1.
 * With kind != 0 the same iteration is right-preconditioned: the Arnoldi
 * process runs on A M^-1 (w = A M^-1 v_j) and the cycle's update is
 * x += M^-1 (sum_i y_i v_i), so the residual, its norm and the stopping test
 * stay those of A x = b. */
static int oracle_gmres_core(int n, const double *a, long lda, const double *b, const double *x0, int max_iter,
                             double tol, int restart, int kind, const double *lu, const int64_t *perm, double *x,
                             int *iterations, double *residual_norm, int *converged)
{
    if (restart <= 0) restart = n < 30 ? n : 30;
    double *v = (double *)calloc((size_t)(restart + 1) * n, sizeof(double));
    double *h = (double *)calloc((size_t)(restart + 1) * restart, sizeof(double));
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    double *wo = (double *)malloc(sizeof(double) * (size_t)n);
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    double *y = (double *)malloc(sizeof(double) * (size_t)(restart + 1));
    double *z = (double *)malloc(sizeof(double) * (size_t)n);
    double *t = (double *)malloc(sizeof(double) * (size_t)n);
    double *dinv = (double *)malloc(sizeof(double) * (size_t)n);
    if (!v || !h || !w || !wo || !r || !y || !z || !t || !dinv) {
        free(v); free(h); free(w); free(wo); free(r); free(y); free(z); free(t); free(dinv);
        return -3;
    }
    if (kind == 1)
        for (int i = 0; i < n; ++i) {
            const double d = a[(size_t)i * lda + i];
            dinv[i] = fabs(d) < DBL_EPSILON ? 1.0 : 1.0 / d;
        }
    for (int i = 0; i < n; ++i) x[i] = x0 ? x0[i] : 0.0;
    int total = 0, rc = 0;
    int cycles = max_iter / restart;
    if (cycles < 1) cycles = 1;
    *converged = 0;
    for (int outer = 0; outer < cycles; ++outer) {
        const double beta = oracle_resid_norm(n, a, lda, b, x, r);
        if (beta < tol) {
            *iterations = total; *residual_norm = beta; *converged = 1;
            goto done;
        }
        for (int i = 0; i < n; ++i) v[i] = r[i] / beta;
        memset(h, 0, sizeof(double) * (size_t)(restart + 1) * restart);
        int j = 0;
        while (j < restart && total < max_iter) {
            const double *vj = v + (size_t)j * n;
            oracle_prec_apply(n, kind, dinv, lu, perm, vj, z);
            for (int i = 0; i < n; ++i) {
                double s = 0.0;
                for (int q = 0; q < n; ++q) s += a[(size_t)i * lda + q] * z[q];
                w[i] = s;
            }
            memcpy(wo, w, sizeof(double) * (size_t)n);
            for (int i = 0; i <= j; ++i) {
                const double *vi = v + (size_t)i * n;
                double d = 0.0;
                for (int q = 0; q < n; ++q) d += w[q] * vi[q];
                h[i * restart + j] = d;
                for (int q = 0; q < n; ++q) wo[q] += -d * vi[q];
            }
            double nn = 0.0;
            for (int q = 0; q < n; ++q) nn += wo[q] * wo[q];
            h[(j + 1) * restart + j] = sqrt(nn);
            if (h[(j + 1) * restart + j] < DBL_EPSILON) { j += 1; break; }
            for (int q = 0; q < n; ++q) v[(size_t)(j + 1) * n + q] = wo[q] / h[(j + 1) * restart + j];
            j += 1;
            total += 1;
        }
        {
            /* the leading (j+1) x j block of h, packed with ld j */
            double *hp = (double *)malloc(sizeof(double) * (size_t)(j + 1) * (j > 0 ? j : 1));
            if (!hp) { rc = -3; goto done; }
            for (int i = 0; i <= j; ++i)
                for (int c = 0; c < j; ++c) hp[i * j + c] = h[i * restart + c];
            rc = j > 0 ? oracle_gmres_lsq(j + 1, j, hp, beta, y) : 0;
            free(hp);
            if (rc) goto done;
        }
        if (kind == 0) {
            for (int i = 0; i < j; ++i)
                for (int q = 0; q < n; ++q) x[q] += y[i] * v[(size_t)i * n + q];
        } else {
            for (int q = 0; q < n; ++q) t[q] = 0.0;
            for (int i = 0; i < j; ++i)
                for (int q = 0; q < n; ++q) t[q] += y[i] * v[(size_t)i * n + q];
            oracle_prec_apply(n, kind, dinv, lu, perm, t, z);
            for (int q = 0; q < n; ++q) x[q] += z[q];
        }
        const double rn = oracle_resid_norm(n, a, lda, b, x, r);
        if (rn < tol) {
            *iterations = total; *residual_norm = rn; *converged = 1;
            goto done;
        }
    }
    *iterations = total;
    *residual_norm = oracle_resid_norm(n, a, lda, b, x, r);
done:
    free(v); free(h); free(w); free(wo); free(r); free(y); free(z); free(t); free(dinv);
    return rc;
}

int oracle_gmres_d(int n, const double *a, long lda, const double *b, const double *x0, int max_iter, double tol,
                   int restart, double *x, int *iterations, double *residual_norm, int *converged)
{
    return oracle_gmres_core(n, a, lda, b, x0, max_iter, tol, restart, 0, NULL, NULL, x, iterations, residual_norm,
                             converged);
}

/* The right-preconditioned restatement of the same iteration (the real
 * counterpart of preconditioners.rs:934-1010 `preconditioned_gmres`, whose
 * body never solves the least-squares problem): kind 0 none, 1 Jacobi,
 * 3 the given LU factors (lu n x n packed, perm). */
int oracle_pgmres_d(int n, const double *a, long lda, const double *b, const double *x0, int max_iter, double tol,
                    int restart, int kind, const double *lu, const int64_t *perm, double *x, int *iterations,
                    double *residual_norm, int *converged)
{
    if (kind != 0 && kind != 1 && kind != 3) return -2;
    if (kind == 3 && (!lu || !perm)) return -2;
    return oracle_gmres_core(n, a, lda, b, x0, max_iter, tol, restart, kind, lu, perm, x, iterations, residual_norm,
                             converged);
}

int oracle_version(void) { return 1; }
