// Restarted GMRES of scirs2-linalg/src/solvers/iterative.rs:317-450 on the GPU:
//
//   scirs_gpu_dgmres  <- gmres  solvers/iterative.rs:317-450 (f64)
//   scirs_gpu_dgmres_precond <- preconditioned_gmres  preconditioners.rs:934-1010
//                     (the same iteration, right-preconditioned: the reference's
//                     body never solves its least-squares problem)
//
// The reference's iteration, step for step: (max_iterations / restart).max(1)
// cycles; per cycle r = b - A x, beta = ||r||, V_0 = r / beta; per inner step
// w = A V_j, h_ij = w . V_i against the unmodified w for i <= j (classical
// Gram-Schmidt), V_{j+1} = (w - sum_i h_ij V_i) / h_{j+1,j}, a break that
// counts no iteration when h_{j+1,j} < eps; then the least-squares problem
// min ||beta e1 - H y|| by the reference's Givens sweep (solve_least_squares_gmres,
// iterative.rs:604-664) and x += sum_i y_i V_i in order.
// Every O(n^2) and O(n) operation runs on the GPU (the Krylov basis stays in
// HBM); the (restart+1) x restart Hessenberg system, a few hundred flops, is
// solved on the host between launches, which is also where the reference's
// control flow reads h_{j+1,j} and the residual norms.
#include "api_internal.hpp"
#include <cfloat>
#include <vector>

using namespace scirs;

namespace {

constexpr int NT = 256;

__device__ double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// y = A x, one wave per row
__global__ __launch_bounds__(NT) void gm_gemv_kernel(int n, const double *__restrict__ A, long lda,
                                                     const double *__restrict__ x, double *__restrict__ y)
{
    const int row = blockIdx.x * (NT / 64) + (threadIdx.x >> 6), l = threadIdx.x & 63;
    if (row >= n) return;
    const double *a = A + (long)row * lda;
    double s = 0.0;
    for (int q = l; q < n; q += 64) s += a[q] * x[q];
    s = wsum(s);
    if (l == 0) y[row] = s;
}

// r = b - A x, one wave per row
__global__ __launch_bounds__(NT) void gm_resid_kernel(int n, const double *__restrict__ A, long lda,
                                                      const double *__restrict__ b, const double *__restrict__ x,
                                                      double *__restrict__ r)
{
    const int row = blockIdx.x * (NT / 64) + (threadIdx.x >> 6), l = threadIdx.x & 63;
    if (row >= n) return;
    const double *a = A + (long)row * lda;
    double s = 0.0;
    for (int q = l; q < n; q += 64) s += a[q] * x[q];
    s = wsum(s);
    if (l == 0) r[row] = b[row] - s;
}

// out[i] = dot(U_i, w) for i < k (U_i = U + i*n); one workgroup per i, fixed order
__global__ __launch_bounds__(1024) void gm_dots_kernel(int n, int k, const double *__restrict__ U,
                                                       const double *__restrict__ w, double *__restrict__ out)
{
    __shared__ double red[16];
    const int i = blockIdx.x;
    if (i >= k) return;
    const double *u = U + (long)i * n;
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += 1024) s += u[q] * w[q];
    s = wsum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < 16; ++q) t += red[q];
        out[i] = t;
    }
}

// dst = (w - sum_{i<k} h_i V_i), the terms subtracted in order i = 0, 1, ...
__global__ __launch_bounds__(NT) void gm_orth_kernel(int n, int k, const double *__restrict__ V,
                                                     const double *__restrict__ h, const double *__restrict__ w,
                                                     double *__restrict__ dst)
{
    const int q = blockIdx.x * NT + threadIdx.x;
    if (q >= n) return;
    double s = w[q];
    for (int i = 0; i < k; ++i) s += -h[i] * V[(long)i * n + q];
    dst[q] = s;
}

// v = v / d (in place, true division)
__global__ __launch_bounds__(NT) void gm_div_kernel(int n, double *__restrict__ v, double d)
{
    const int q = blockIdx.x * NT + threadIdx.x;
    if (q < n) v[q] = v[q] / d;
}

// dst = src / d
__global__ __launch_bounds__(NT) void gm_div_copy_kernel(int n, const double *__restrict__ src, double d,
                                                         double *__restrict__ dst)
{
    const int q = blockIdx.x * NT + threadIdx.x;
    if (q < n) dst[q] = src[q] / d;
}

// x += sum_{i<k} y_i V_i, added in order
__global__ __launch_bounds__(NT) void gm_combine_kernel(int n, int k, const double *__restrict__ V,
                                                        const double *__restrict__ y, double *__restrict__ x)
{
    const int q = blockIdx.x * NT + threadIdx.x;
    if (q >= n) return;
    double s = x[q];
    for (int i = 0; i < k; ++i) s += y[i] * V[(long)i * n + q];
    x[q] = s;
}

// z = dinv * v (the Jacobi preconditioner)
__global__ __launch_bounds__(NT) void gm_scale_vec_kernel(int n, const double *__restrict__ dinv,
                                                          const double *__restrict__ v, double *__restrict__ z)
{
    const int q = blockIdx.x * NT + threadIdx.x;
    if (q < n) z[q] = dinv[q] * v[q];
}

// dinv[i] = 1 / a_ii, or 1 where |a_ii| < eps (preconditioners.rs:210-229)
__global__ __launch_bounds__(NT) void gm_jacobi_kernel(int n, const double *__restrict__ A, long lda,
                                                       double *__restrict__ dinv)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    const double d = A[(long)i * lda + i];
    dinv[i] = fabs(d) < DBL_EPSILON ? 1.0 : 1.0 / d;
}

// dst = sum_{i<k} y_i V_i, added in order from 0
__global__ __launch_bounds__(NT) void gm_basis_combine_kernel(int n, int k, const double *__restrict__ V,
                                                              const double *__restrict__ y, double *__restrict__ dst)
{
    const int q = blockIdx.x * NT + threadIdx.x;
    if (q >= n) return;
    double s = 0.0;
    for (int i = 0; i < k; ++i) s += y[i] * V[(long)i * n + q];
    dst[q] = s;
}

// x += z
__global__ __launch_bounds__(NT) void gm_add_kernel(int n, const double *__restrict__ z, double *__restrict__ x)
{
    const int q = blockIdx.x * NT + threadIdx.x;
    if (q < n) x[q] += z[q];
}

// iterative.rs:604-664: Givens sweep over the m x nc Hessenberg block, then
// back substitution; returns i+1 when |r_ii| < eps.
int gmres_lsq(int m, int nc, std::vector<double> hc, double beta, std::vector<double> &y)
{
    std::vector<double> g(m, 0.0);
    g[0] = beta;
    for (int j = 0; j < nc; ++j)
        for (int i = j + 1; i < m; ++i) {
            if (fabs(hc[(size_t)i * nc + j]) > DBL_EPSILON) {
                const double a = hc[(size_t)j * nc + j], b = hc[(size_t)i * nc + j];
                const double r = sqrt(a * a + b * b);
                const double c = a / r, s = b / r;
                for (int k = j; k < nc; ++k) {
                    const double hjk = hc[(size_t)j * nc + k], hik = hc[(size_t)i * nc + k];
                    hc[(size_t)j * nc + k] = c * hjk + s * hik;
                    hc[(size_t)i * nc + k] = -s * hjk + c * hik;
                }
                const double gj = g[j], gi = g[i];
                g[j] = c * gj + s * gi;
                g[i] = -s * gj + c * gi;
            }
        }
    y.assign(nc, 0.0);
    for (int i = nc - 1; i >= 0; --i) {
        double sum = g[i];
        for (int j = i + 1; j < nc; ++j) sum -= hc[(size_t)i * nc + j] * y[j];
        if (fabs(hc[(size_t)i * nc + i]) < DBL_EPSILON) return i + 1;
        y[i] = sum / hc[(size_t)i * nc + i];
    }
    return 0;
}

unsigned blocks(int n, int per) { return (unsigned)((n + per - 1) / per); }

// Right preconditioner M^-1 on the device: none, Jacobi (dinv) or LU factors
// in FP64 (lu64, n x n packed, P A = L U with perm) applied by permutation and
// the two chained triangular solves.
struct Precond {
    int kind = SCIRS_PRECOND_NONE;
    const double *dinv = nullptr;
    const double *lu64 = nullptr;
    const int64_t *perm = nullptr;
};

int precond_apply(const Precond &m, int n, const double *v, double *z, hipStream_t s)
{
    if (m.kind == SCIRS_PRECOND_DIAGONAL) {
        gm_scale_vec_kernel<<<blocks(n, NT), NT, 0, s>>>(n, m.dinv, v, z);
        SCIRS_HIP_CHECK(hipGetLastError());
        return SCIRS_OK;
    }
    if (int rc = scirs_permute_rows<double>(n, 1, m.perm, v, 1, z, 1, s)) return rc;
    return scirs::lu_substitute<double>(n, 1, m.lu64, n, z, 1, s);
}

// The iteration on device-resident A (lda), b and x (x holds x0 on entry, the
// solution on return), on stream s; the Krylov basis and small vectors live in
// the library's scratch.  With a preconditioner the Arnoldi process runs on
// A M^-1 and each cycle ends with x += M^-1 (V y): right preconditioning, so
// the residual and the stopping test stay those of A x = b.  Returns 0,
// k > 0 (least squares singular at row k) or an SCIRS_ERR_* code.
int gmres_device(int n, const double *dA, long lda, const double *db, double *dx, int max_iter, double tol,
                 int restart, int *iterations, double *residual_norm, int *converged, hipStream_t s,
                 scirs_solve_info *info, const Precond &pm = Precond())
{
    Context &c = ctx();
    int rc;
    const size_t nn = (size_t)n;
    const size_t wbytes = ((size_t)(restart + 1) * nn + 3 * nn + 2 * (size_t)(restart + 2) + 64) * sizeof(double);
    if ((rc = c.ensure(&c.scratch_b, &c.scratch_b_bytes, wbytes))) return rc;
    double *dV = (double *)c.scratch_b, *dr = dV + (size_t)(restart + 1) * nn, *dw = dr + nn, *dz = dw + nn,
           *dh = dz + nn, *dy = dh + (restart + 2), *dsc = dy + (restart + 2);
    const bool prec = pm.kind != SCIRS_PRECOND_NONE;

    // ||b - A x|| into the host; r stays in dr
    auto resid_norm = [&](double *out) -> int {
        gm_resid_kernel<<<blocks(n, NT / 64), NT, 0, s>>>(n, dA, lda, db, dx, dr);
        gm_dots_kernel<<<1, 1024, 0, s>>>(n, 1, dr, dr, dsc);
        SCIRS_HIP_CHECK(hipGetLastError());
        double v;
        SCIRS_HIP_CHECK(hipMemcpyAsync(&v, dsc, sizeof(double), hipMemcpyDeviceToHost, s));
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
        *out = sqrt(v);
        return 0;
    };
    auto done = [&](int total, double rn, int conv) -> int {
        *iterations = total;
        *residual_norm = rn;
        *converged = conv;
        if (info) info->gmres_iters = total;
        return 0;
    };

    const int cycles = std::max(max_iter / restart, 1);
    int total = 0;
    std::vector<double> H((size_t)(restart + 1) * restart), hcol(restart + 2), y;
    for (int outer = 0; outer < cycles; ++outer) {
        double beta;
        if ((rc = resid_norm(&beta))) return rc;
        if (beta < tol) return done(total, beta, 1);
        gm_div_copy_kernel<<<blocks(n, NT), NT, 0, s>>>(n, dr, beta, dV);
        std::fill(H.begin(), H.end(), 0.0);
        int j = 0;
        while (j < restart && total < max_iter) {
            double *vj = dV + (size_t)j * nn, *vn = vj + nn;
            if (prec) {
                // w = A M^-1 v_j
                if ((rc = precond_apply(pm, n, vj, dz, s))) return rc;
                gm_gemv_kernel<<<blocks(n, NT / 64), NT, 0, s>>>(n, dA, lda, dz, dw);
            } else {
                gm_gemv_kernel<<<blocks(n, NT / 64), NT, 0, s>>>(n, dA, lda, vj, dw);
            }
            gm_dots_kernel<<<j + 1, 1024, 0, s>>>(n, j + 1, dV, dw, dh);
            gm_orth_kernel<<<blocks(n, NT), NT, 0, s>>>(n, j + 1, dV, dh, dw, vn);
            gm_dots_kernel<<<1, 1024, 0, s>>>(n, 1, vn, vn, dh + j + 1);
            SCIRS_HIP_CHECK(hipGetLastError());
            SCIRS_HIP_CHECK(hipMemcpyAsync(hcol.data(), dh, (j + 2) * sizeof(double), hipMemcpyDeviceToHost, s));
            SCIRS_HIP_CHECK(hipStreamSynchronize(s));
            hcol[j + 1] = sqrt(hcol[j + 1]);
            for (int i = 0; i <= j + 1; ++i) H[(size_t)i * restart + j] = hcol[i];
            if (hcol[j + 1] < DBL_EPSILON) {
                j += 1;
                break;
            }
            gm_div_kernel<<<blocks(n, NT), NT, 0, s>>>(n, vn, hcol[j + 1]);
            j += 1;
            total += 1;
        }
        std::vector<double> hp((size_t)(j + 1) * std::max(j, 1));
        for (int i = 0; i <= j; ++i)
            for (int cc = 0; cc < j; ++cc) hp[(size_t)i * j + cc] = H[(size_t)i * restart + cc];
        if (j > 0) {
            const int st = gmres_lsq(j + 1, j, hp, beta, y);
            if (st) {
                scirs_set_error("dgmres: least-squares solve singular at row %d", st);
                if (info) info->gmres_iters = total;
                return st;
            }
            SCIRS_HIP_CHECK(hipMemcpyAsync(dy, y.data(), j * sizeof(double), hipMemcpyHostToDevice, s));
            if (prec) {
                // x += M^-1 (V y)
                gm_basis_combine_kernel<<<blocks(n, NT), NT, 0, s>>>(n, j, dV, dy, dw);
                if ((rc = precond_apply(pm, n, dw, dz, s))) return rc;
                gm_add_kernel<<<blocks(n, NT), NT, 0, s>>>(n, dz, dx);
            } else {
                gm_combine_kernel<<<blocks(n, NT), NT, 0, s>>>(n, j, dV, dy, dx);
            }
            SCIRS_HIP_CHECK(hipGetLastError());
        }
        double rn;
        if ((rc = resid_norm(&rn))) return rc;
        if (rn < tol) return done(total, rn, 1);
    }
    double rn;
    if ((rc = resid_norm(&rn))) return rc;
    return done(total, rn, 0);
}

} // namespace

extern "C" int scirs_gpu_dgmres(int n, const double *a, int64_t lda, const double *b, const double *x0, int max_iter,
                                double tol, int restart, double *x, int *iterations, double *residual_norm,
                                int *converged, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    if (n <= 0 || max_iter < 0 || !a || !b || !x || lda < n || !iterations || !residual_norm || !converged) {
        scirs_set_error("dgmres: invalid arguments (n=%d max_iter=%d)", n, max_iter);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    if (restart <= 0) restart = std::min(n, 30);
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    Context &c = ctx();
    int rc = c.init(0);
    if (rc) return finish(info, rc, t0);
    hipStream_t s = c.stream;
    const size_t nn = (size_t)n;
    if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, (nn * nn + 2 * nn) * sizeof(double))))
        return finish(info, rc, t0);
    double *dA = (double *)c.scratch_a, *db = dA + nn * nn, *dx = db + nn;
    if ((rc = scirs::copy_h2d(dA, nn * sizeof(double), a, lda * sizeof(double), nn * sizeof(double), n, s)))
        return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipMemcpyAsync(db, b, nn * sizeof(double), hipMemcpyHostToDevice, s));
    if (x0) SCIRS_HIP_CHECK(hipMemcpyAsync(dx, x0, nn * sizeof(double), hipMemcpyHostToDevice, s));
    else SCIRS_HIP_CHECK(hipMemsetAsync(dx, 0, nn * sizeof(double), s));
    rc = gmres_device(n, dA, n, db, dx, max_iter, tol, restart, iterations, residual_norm, converged, s, info);
    if (rc) return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipMemcpyAsync(x, dx, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    return finish(info, SCIRS_OK, t0);
}

extern "C" int scirs_gpu_dgmres_dev(int n, const double *dA, int64_t lda, const double *dB, double *dX, int use_x0,
                                    int max_iter, double tol, int restart, int *iterations, double *residual_norm,
                                    int *converged, void *stream, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    if (n <= 0 || max_iter < 0 || !dA || !dB || !dX || lda < n || !iterations || !residual_norm || !converged) {
        scirs_set_error("dgmres_dev: invalid arguments (n=%d max_iter=%d)", n, max_iter);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    if (restart <= 0) restart = std::min(n, 30);
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    int rc = ctx().init(0);
    if (rc) return finish(info, rc, t0);
    hipStream_t s = (hipStream_t)stream;   // NULL: the default stream, ordered after the caller's work on it
    if (!use_x0) SCIRS_HIP_CHECK(hipMemsetAsync(dX, 0, (size_t)n * sizeof(double), s));
    rc = gmres_device(n, dA, lda, dB, dX, max_iter, tol, restart, iterations, residual_norm, converged, s, info);
    return finish(info, rc, t0);
}

namespace {
// The mixed-precision LU of A as the preconditioner: FP32 tournament-pivoted
// panel and FP32-accurate split-BF16 update (the factors of the mixed solve),
// upcast to FP64 so that M^-1 is applied without FP32 rounding.  Returns 0,
// k > 0 when the FP32 factorisation breaks down at column k, or an error.
int lu_preconditioner(int n, const double *dA, long lda, double *lu64, int64_t *perm, hipStream_t s)
{
    Context &c = ctx();
    int rc;
    const long ld32 = ((long)n + 63) / 64 * 64;
    if ((rc = c.ensure(&c.scratch_ir, &c.scratch_ir_bytes, (size_t)n * ld32 * sizeof(float)))) return rc;
    float *a32 = (float *)c.scratch_ir;
    if ((rc = scirs_d2s(n, n, dA, lda, a32, ld32, s))) return rc;
    if ((rc = c.reset_flags(s))) return rc;
    const bool tslu = c.work.tslu != 0;
    const int kb = tslu ? c.kb_f32 : c.lu_kb<float>(n);
    const int saved = c.work.lowp_mode;
    c.work.lowp_mode = SCIRS_LOWP_BF16X6;
    c.work.tslu_active = tslu;
    rc = scirs_getrf_device<float>(n, n, a32, ld32, kb, c.work, s);
    c.work.tslu_active = 0;
    c.work.lowp_mode = saved;
    if (rc) return rc;
    if ((rc = scirs_perm_from_pairs(n, n, kb, c.work, perm, s))) return rc;
    int st = 0;
    if ((rc = c.read_flags(s, &st, nullptr))) return rc;
    if (st) return st;
    return scirs_s2d(n, n, a32, ld32, lu64, n, s);
}

// Builds the device preconditioner of kind `precond` for device-resident A;
// for SCIRS_PRECOND_LU_GIVEN, dLU / dPerm are the caller's device factors.
int make_precond(int n, const double *dA, long lda, int precond, const double *dLU, const int64_t *dPerm,
                 hipStream_t s, Precond &pm)
{
    Context &c = ctx();
    int rc;
    pm.kind = precond;
    if (precond == SCIRS_PRECOND_NONE) return SCIRS_OK;
    if (precond == SCIRS_PRECOND_DIAGONAL) {
        if ((rc = c.ensure(&c.scratch_perm, &c.scratch_perm_bytes, (size_t)n * sizeof(double)))) return rc;
        double *dinv = (double *)c.scratch_perm;
        gm_jacobi_kernel<<<blocks(n, NT), NT, 0, s>>>(n, dA, lda, dinv);
        SCIRS_HIP_CHECK(hipGetLastError());
        pm.dinv = dinv;
        return SCIRS_OK;
    }
    if (precond == SCIRS_PRECOND_LU_GIVEN) {
        // the chained triangular solves that apply it skip their work while the
        // breakdown flag holds a previous call's value
        if ((rc = c.reset_flags(s))) return rc;
        pm.lu64 = dLU;
        pm.perm = dPerm;
        return SCIRS_OK;
    }
    // SCIRS_PRECOND_LU
    if ((rc = c.ensure(&c.scratch_lu64, &c.scratch_lu64_bytes, (size_t)n * n * sizeof(double)))) return rc;
    if ((rc = c.ensure(&c.scratch_perm, &c.scratch_perm_bytes, (size_t)n * sizeof(int64_t)))) return rc;
    double *lu64 = (double *)c.scratch_lu64;
    int64_t *perm = (int64_t *)c.scratch_perm;
    if ((rc = lu_preconditioner(n, dA, lda, lu64, perm, s))) return rc;
    pm.lu64 = lu64;
    pm.perm = perm;
    return SCIRS_OK;
}

// After GMRES with an LU preconditioner: a chained triangular solve that timed
// out (its wait is bounded) is reported, not returned as a solution.
int check_precond_flags(const Precond &pm, hipStream_t s)
{
    if (pm.kind != SCIRS_PRECOND_LU && pm.kind != SCIRS_PRECOND_LU_GIVEN) return SCIRS_OK;
    int st = 0;
    return ctx().read_flags(s, &st, nullptr);
}

bool precond_kind_ok(int k)
{
    return k == SCIRS_PRECOND_NONE || k == SCIRS_PRECOND_DIAGONAL || k == SCIRS_PRECOND_LU ||
           k == SCIRS_PRECOND_LU_GIVEN;
}
} // namespace

extern "C" int scirs_gpu_dgmres_precond(int n, const double *a, int64_t lda, const double *b, const double *x0,
                                        int max_iter, double tol, int restart, int precond, const double *lu,
                                        const int64_t *perm, double *x, int *iterations, double *residual_norm,
                                        int *converged, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    if (n <= 0 || max_iter < 0 || !a || !b || !x || lda < n || !iterations || !residual_norm || !converged ||
        !precond_kind_ok(precond) || (precond == SCIRS_PRECOND_LU_GIVEN && (!lu || !perm))) {
        scirs_set_error("dgmres_precond: invalid arguments (n=%d max_iter=%d precond=%d; the given-LU "
                        "preconditioner needs lu and perm)", n, max_iter, precond);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    if (precond == SCIRS_PRECOND_LU_GIVEN)
        for (int i = 0; i < n; ++i)
            if (perm[i] < 0 || perm[i] >= n) {
                scirs_set_error("dgmres_precond: perm[%d] = %lld is not a row index", i, (long long)perm[i]);
                return finish(info, SCIRS_ERR_ARG, t0);
            }
    if (restart <= 0) restart = std::min(n, 30);
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    Context &c = ctx();
    int rc = c.init(0);
    if (rc) return finish(info, rc, t0);
    hipStream_t s = c.stream;
    const size_t nn = (size_t)n;
    const size_t extra = precond == SCIRS_PRECOND_LU_GIVEN ? nn * nn + nn : 0;   // given factors and perm
    if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, (nn * nn + 2 * nn + extra) * sizeof(double))))
        return finish(info, rc, t0);
    double *dA = (double *)c.scratch_a, *db = dA + nn * nn, *dx = db + nn;
    double *dLU = extra ? dx + nn : nullptr;
    int64_t *dP = extra ? (int64_t *)(dLU + nn * nn) : nullptr;
    if ((rc = scirs::copy_h2d(dA, nn * sizeof(double), a, lda * sizeof(double), nn * sizeof(double), n, s)))
        return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipMemcpyAsync(db, b, nn * sizeof(double), hipMemcpyHostToDevice, s));
    if (extra) {
        if ((rc = scirs::copy_h2d_flat(dLU, lu, nn * nn * sizeof(double), s))) return finish(info, rc, t0);
        SCIRS_HIP_CHECK(hipMemcpyAsync(dP, perm, nn * sizeof(int64_t), hipMemcpyHostToDevice, s));
    }
    if (x0) SCIRS_HIP_CHECK(hipMemcpyAsync(dx, x0, nn * sizeof(double), hipMemcpyHostToDevice, s));
    else SCIRS_HIP_CHECK(hipMemsetAsync(dx, 0, nn * sizeof(double), s));
    Precond pm;
    if ((rc = make_precond(n, dA, n, precond, dLU, dP, s, pm))) return finish(info, rc, t0);
    rc = gmres_device(n, dA, n, db, dx, max_iter, tol, restart, iterations, residual_norm, converged, s, info, pm);
    if (rc || (rc = check_precond_flags(pm, s))) return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipMemcpyAsync(x, dx, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    return finish(info, SCIRS_OK, t0);
}

extern "C" int scirs_gpu_dgmres_precond_dev(int n, const double *dA, int64_t lda, const double *dB, double *dX,
                                            int use_x0, int max_iter, double tol, int restart, int precond,
                                            const double *dLU, const int64_t *dPerm, int *iterations,
                                            double *residual_norm, int *converged, void *stream,
                                            scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    if (n <= 0 || max_iter < 0 || !dA || !dB || !dX || lda < n || !iterations || !residual_norm || !converged ||
        !precond_kind_ok(precond) || (precond == SCIRS_PRECOND_LU_GIVEN && (!dLU || !dPerm))) {
        scirs_set_error("dgmres_precond_dev: invalid arguments (n=%d max_iter=%d precond=%d; the given-LU "
                        "preconditioner needs dLU and dPerm)", n, max_iter, precond);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    if (restart <= 0) restart = std::min(n, 30);
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    int rc = ctx().init(0);
    if (rc) return finish(info, rc, t0);
    hipStream_t s = (hipStream_t)stream;   // NULL: the default stream, ordered after the caller's work on it
    if (!use_x0) SCIRS_HIP_CHECK(hipMemsetAsync(dX, 0, (size_t)n * sizeof(double), s));
    Precond pm;
    if ((rc = make_precond(n, dA, lda, precond, dLU, dPerm, s, pm))) return finish(info, rc, t0);
    rc = gmres_device(n, dA, lda, dB, dX, max_iter, tol, restart, iterations, residual_norm, converged, s, info, pm);
    if (!rc) rc = check_precond_flags(pm, s);
    return finish(info, rc, t0);
}
