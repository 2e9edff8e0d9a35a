// Restarted GMRES of scirs2-linalg/src/solvers/iterative.rs:317-450 on the GPU:
//
//   scirs_gpu_dgmres  <- gmres  solvers/iterative.rs:317-450 (f64)
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

// The iteration on device-resident A (lda), b and x (x holds x0 on entry, the
// solution on return), on stream s; the Krylov basis and small vectors live in
// the library's scratch.  Returns 0, k > 0 (least squares singular at row k)
// or an SCIRS_ERR_* code.
int gmres_device(int n, const double *dA, long lda, const double *db, double *dx, int max_iter, double tol,
                 int restart, int *iterations, double *residual_norm, int *converged, hipStream_t s,
                 scirs_solve_info *info)
{
    Context &c = ctx();
    int rc;
    const size_t nn = (size_t)n;
    const size_t wbytes = ((size_t)(restart + 1) * nn + 2 * nn + 2 * (size_t)(restart + 2) + 64) * sizeof(double);
    if ((rc = c.ensure(&c.scratch_b, &c.scratch_b_bytes, wbytes))) return rc;
    double *dV = (double *)c.scratch_b, *dr = dV + (size_t)(restart + 1) * nn, *dw = dr + nn, *dh = dw + nn,
           *dy = dh + (restart + 2), *dsc = dy + (restart + 2);

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
            gm_gemv_kernel<<<blocks(n, NT / 64), NT, 0, s>>>(n, dA, lda, vj, dw);
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
            gm_combine_kernel<<<blocks(n, NT), NT, 0, s>>>(n, j, dV, dy, dx);
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
    SCIRS_HIP_CHECK(hipMemcpy2DAsync(dA, nn * sizeof(double), a, lda * sizeof(double), nn * sizeof(double), n,
                                     hipMemcpyHostToDevice, s));
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
