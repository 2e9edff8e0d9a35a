// Mixed-precision solve: factor once in low precision, refine in FP64.
//
// Counterpart of scirs2-linalg/src/mixed_precision/adaptive.rs:282-378
// (`iterative_refinement_solve`), which re-factors A in the working precision
// on every step and stops on an absolute residual.  Here:
//   1. A32 = fp32(A) fused with the row sums of |A| (||A||_inf);
//   2. LU of A32: FP32 panel, trailing update in the requested arithmetic
//      (BF16 / FP16-scaled / split-BF16 / FP32 matrix cores, FP32 accumulation);
//   3. refinement  R = B - A X (FP64, GPU), D = U^-1 L^-1 P R (FP32),
//      X += D, stopping when the normwise backward error
//      max_j ||R_j||_inf / (||A||_inf ||X_j||_inf + ||B_j||_inf) <= tol;
//   4. when a step fails to halve the backward error, restarted GMRES (FP64
//      Arnoldi with classical Gram-Schmidt applied twice, right-preconditioned
//      by the low-precision factors) solves the correction equation A D = R
//      (Carson & Higham, SISC 40(2), 2018, "GMRES-IR"); the Givens least-squares
//      update is the one of scirs2-linalg/src/solvers/iterative.rs:604-664;
//   5. when GMRES-IR also stalls, or the low-precision LU breaks down, FP64 LU
//      solves the system (the breakdown is then decided the FP64 way).
// FP64 is used on the GPU only for the residual, the update of X and the
// stopping test (plus the Arnoldi vectors of GMRES).
#include "api_internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

using scirs::Context;
using scirs::ctx;

constexpr int NT = 256;

// y = A x, FP64 GEMV; one wave per 4 rows.
__global__ __launch_bounds__(NT) void gemv_kernel(int n, const double *A, long lda, const double *x, double *y)
{
    const int lane = threadIdx.x & 63;
    const int r0 = ((blockIdx.x * NT + threadIdx.x) >> 6) * 4;
    if (r0 >= n) return;
    const int nr = min(4, n - r0);
    double acc[4] = {0, 0, 0, 0};
    for (int k = lane; k < n; k += 64) {
        const double xv = x[k];
#pragma unroll
        for (int a = 0; a < 4; ++a)
            if (a < nr) acc[a] = fma(A[(long)(r0 + a) * lda + k], xv, acc[a]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) acc[a] = wave_sum(acc[a]);
    if (lane == 0)
        for (int a = 0; a < nr; ++a) y[r0 + a] = acc[a];
}

// h[i] = dot(V_i, w) for i < k (V: k vectors of length n, row i contiguous)
__global__ __launch_bounds__(NT) void multidot_kernel(int n, int k, const double *V, const double *w, double *h)
{
    const int i = blockIdx.x;
    if (i >= k) return;
    const double *v = V + (long)i * n;
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += NT) s = fma(v[q], w[q], s);
    s = wave_sum(s);
    __shared__ double red[NT / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int q = 0; q < NT / 64; ++q) r += red[q];
        h[i] = r;
    }
}

// w -= sum_i h[i] V_i
__global__ void mgs_update_kernel(int n, int k, const double *V, const double *h, double *w)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double s = w[q];
    for (int i = 0; i < k; ++i) s -= h[i] * V[(long)i * n + q];
    w[q] = s;
}

// dst = src * scale
__global__ void scale_copy_kernel(int n, const double *src, double scale, double *dst)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = src[q] * scale;
}

// d32[i] = fp32(r[perm[i] * ld + col] * scale)   (row permutation + scaling + rounding)
__global__ void permute_to_f32_kernel(int n, const int64_t *perm, const double *r, long ld, int col, double scale,
                                      float *d32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d32[i] = (float)(r[perm[i] * ld + col] * scale);
}

// x[i * ld + col] += scale * d32[i]   (or =, when assign)
__global__ void axpy_f32_kernel(int n, const float *d32, double scale, double *x, long ld, int col, int assign)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = scale * (double)d32[i];
    if (assign) x[(long)i * ld + col] = v; else x[(long)i * ld + col] += v;
}

// z = fp64(d32) * scale
__global__ void f32_to_f64_scaled_kernel(int n, const float *d32, double scale, double *z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = scale * (double)d32[i];
}

// x[i * ld + col] += sum_j y[j] * Z_j[i]
__global__ void combine_kernel(int n, int k, const double *Z, const double *y, double *x, long ld, int col)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += y[j] * Z[(long)j * n + i];
    x[(long)i * ld + col] += s;
}

__global__ void zero_ld_kernel(int rows, int cols, double *p, long ld)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (long)rows * cols) p[(i / cols) * ld + i % cols] = 0.0;
}

inline unsigned nb(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// power of two closest above |v| (1 for v == 0), so scaling is exact
double pow2_above(double v)
{
    if (!(v > 0) || !std::isfinite(v)) return 1.0;
    int e;
    std::frexp(v, &e);
    return std::ldexp(1.0, e);
}

struct IrState {
    int n, nrhs;
    const double *A; long lda;
    const double *B; long ldb;
    double *X; long ldx;
    float *LU; long ld32; int64_t *perm;
    double *R;           // n x nrhs residual
    float *d32;          // n
    double *norms;       // device: [0..nrhs) r, [64..) x, [128..) b, [192] A
    double *rowsum;      // n
    hipStream_t s;
    double normA;
    std::vector<double> normB;
    int kb;
    const float *dinv_l = nullptr, *dinv_u = nullptr;   // inverted 64 x 64 diagonal blocks of L and U
    const float *dmt_l = nullptr, *dmt_u = nullptr;     // those times each block row's last tile
};

// d32 <- U^-1 L^-1 d32 with the FP32 factors
int lu_solve_f32(IrState &st, float *d32)
{
    return scirs::lu_substitute<float>(st.n, 1, st.LU, st.ld32, d32, 1, st.s, st.dinv_l, st.dinv_u, st.dmt_l,
                                       st.dmt_u);
}

// apply z = (LU)^-1 P v in FP32 for one vector v (FP64), result FP64; scale keeps FP32 in range
int precond_apply(IrState &st, const double *v, double vnorm, double *z)
{
    Context &c = ctx();
    const double sc = pow2_above(vnorm);
    hipLaunchKernelGGL(permute_to_f32_kernel, dim3(nb(st.n)), dim3(256), 0, st.s, st.n, st.perm, v, 1L, 0, 1.0 / sc,
                       st.d32);
    int rc = lu_solve_f32(st, st.d32);
    if (rc) return rc;
    hipLaunchKernelGGL(f32_to_f64_scaled_kernel, dim3(nb(st.n)), dim3(256), 0, st.s, st.n, st.d32, sc, z);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

// eta_j for every column; returns the worst, fills per-column values
int backward_errors(IrState &st, std::vector<double> &eta, std::vector<double> &rnorm)
{
    int rc;
    if ((rc = scirs_residual(st.n, st.nrhs, st.A, st.lda, st.X, st.ldx, st.B, st.ldb, st.R, st.nrhs, st.s))) return rc;
    if ((rc = scirs_colmax_d(st.n, st.nrhs, st.R, st.nrhs, st.norms, st.s))) return rc;
    if ((rc = scirs_colmax_d(st.n, st.nrhs, st.X, st.ldx, st.norms + 64, st.s))) return rc;
    double h[128];
    SCIRS_HIP_CHECK(hipMemcpyAsync(h, st.norms, 128 * sizeof(double), hipMemcpyDeviceToHost, st.s));
    SCIRS_HIP_CHECK(hipStreamSynchronize(st.s));
    eta.assign(st.nrhs, 0.0);
    rnorm.assign(st.nrhs, 0.0);
    for (int j = 0; j < st.nrhs; ++j) {
        const double den = st.normA * h[64 + j] + st.normB[j];
        rnorm[j] = h[j];
        eta[j] = den > 0 ? h[j] / den : (h[j] > 0 ? INFINITY : 0.0);
    }
    return SCIRS_OK;
}

// one LU-IR correction for column j: X_j += (LU)^-1 P R_j
int ir_step(IrState &st, int j, double rn)
{
    Context &c = ctx();
    const double sc = pow2_above(rn);
    hipLaunchKernelGGL(permute_to_f32_kernel, dim3(nb(st.n)), dim3(256), 0, st.s, st.n, st.perm, st.R, (long)st.nrhs, j,
                       1.0 / sc, st.d32);
    int rc = lu_solve_f32(st, st.d32);
    if (rc) return rc;
    hipLaunchKernelGGL(axpy_f32_kernel, dim3(nb(st.n)), dim3(256), 0, st.s, st.n, st.d32, sc, st.X, st.ldx, j, 0);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

// Restarted right-preconditioned GMRES on A d = R_j (FP64 Arnoldi), X_j += d.
// Returns the number of inner iterations used in *iters.
// Restarted GMRES on A d = r_j, right-preconditioned by the low-precision LU.
// *stalled is set when a whole restart cycle fails to halve the true residual:
// the preconditioner is then too weak for this system, and the caller stops
// spending iterations on it.
int gmres_correction(IrState &st, int j, int restart, int max_inner, double rtol, double *V, double *Z, double *w,
                     double *hdev, int *iters, bool *stalled)
{
    *stalled = false;
    const int n = st.n;
    // r0 = R_j (copy of column j)
    std::vector<double> H((size_t)(restart + 1) * restart, 0.0), cs(restart), sn(restart), g(restart + 1);
    int rc;
    // V_0 <- R[:, j]
    if ((rc = scirs_d2d_copy(n, 1, st.R + j, st.nrhs, V, 1, st.s))) return rc;
    double beta;
    hipLaunchKernelGGL(multidot_kernel, dim3(1), dim3(NT), 0, st.s, n, 1, V, V, hdev);
    SCIRS_HIP_CHECK(hipMemcpyAsync(&beta, hdev, sizeof(double), hipMemcpyDeviceToHost, st.s));
    SCIRS_HIP_CHECK(hipStreamSynchronize(st.s));
    beta = std::sqrt(beta);
    *iters = 0;
    if (!(beta > 0)) return SCIRS_OK;
    const double target = rtol * beta;
    int total = 0;
    while (total < max_inner) {
        hipLaunchKernelGGL(scale_copy_kernel, dim3(nb(n)), dim3(256), 0, st.s, n, V, 1.0 / beta, V);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        double resid = beta;
        for (; k < restart && total < max_inner; ++k, ++total) {
            double *vk = V + (long)k * n, *vk1 = V + (long)(k + 1) * n, *zk = Z + (long)k * n;
            // z_k = M^-1 v_k ;  w = A z_k
            if ((rc = precond_apply(st, vk, 1.0, zk))) return rc;
            hipLaunchKernelGGL(gemv_kernel, dim3(nb((long)((n + 3) / 4) * 64, NT)), dim3(NT), 0, st.s, n, st.A, st.lda,
                               zk, w);
            // classical Gram-Schmidt, twice
            std::vector<double> hcol(k + 2, 0.0), tmp(k + 1);
            for (int pass = 0; pass < 2; ++pass) {
                hipLaunchKernelGGL(multidot_kernel, dim3(k + 1), dim3(NT), 0, st.s, n, k + 1, V, w, hdev);
                hipLaunchKernelGGL(mgs_update_kernel, dim3(nb(n)), dim3(256), 0, st.s, n, k + 1, V, hdev, w);
                SCIRS_HIP_CHECK(hipMemcpyAsync(tmp.data(), hdev, (k + 1) * sizeof(double), hipMemcpyDeviceToHost,
                                               st.s));
                SCIRS_HIP_CHECK(hipStreamSynchronize(st.s));
                for (int i = 0; i <= k; ++i) hcol[i] += tmp[i];
            }
            double hn;
            hipLaunchKernelGGL(multidot_kernel, dim3(1), dim3(NT), 0, st.s, n, 1, w, w, hdev);
            SCIRS_HIP_CHECK(hipMemcpyAsync(&hn, hdev, sizeof(double), hipMemcpyDeviceToHost, st.s));
            SCIRS_HIP_CHECK(hipStreamSynchronize(st.s));
            hn = std::sqrt(hn);
            hcol[k + 1] = hn;
            if (hn > 0) hipLaunchKernelGGL(scale_copy_kernel, dim3(nb(n)), dim3(256), 0, st.s, n, w, 1.0 / hn, vk1);
            // Givens rotations (as solvers/iterative.rs:604-664 solves the small LS problem)
            for (int i = 0; i < k; ++i) {
                const double a = hcol[i], b = hcol[i + 1];
                hcol[i] = cs[i] * a + sn[i] * b;
                hcol[i + 1] = -sn[i] * a + cs[i] * b;
            }
            const double a = hcol[k], b = hcol[k + 1];
            const double r = std::hypot(a, b);
            cs[k] = r > 0 ? a / r : 1.0;
            sn[k] = r > 0 ? b / r : 0.0;
            hcol[k] = r;
            hcol[k + 1] = 0.0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            for (int i = 0; i <= k; ++i) H[(size_t)i * restart + k] = hcol[i];
            resid = std::fabs(g[k + 1]);
            if (resid <= target || !(hn > 0)) { ++k; ++total; break; }
        }
        // back substitution for y, then X_j += Z y
        std::vector<double> y(k, 0.0);
        for (int i = k - 1; i >= 0; --i) {
            double sacc = g[i];
            for (int q = i + 1; q < k; ++q) sacc -= H[(size_t)i * restart + q] * y[q];
            y[i] = sacc / H[(size_t)i * restart + i];
        }
        SCIRS_HIP_CHECK(hipMemcpyAsync(hdev, y.data(), k * sizeof(double), hipMemcpyHostToDevice, st.s));
        hipLaunchKernelGGL(combine_kernel, dim3(nb(n)), dim3(256), 0, st.s, n, k, Z, hdev, st.X, st.ldx, j);
        SCIRS_HIP_CHECK(hipGetLastError());
        if (resid <= target || total >= max_inner) break;
        // restart from the new residual of this column
        if ((rc = scirs_residual(n, 1, st.A, st.lda, st.X + j, st.ldx, st.B + j, st.ldb, V, 1, st.s))) return rc;
        hipLaunchKernelGGL(multidot_kernel, dim3(1), dim3(NT), 0, st.s, n, 1, V, V, hdev);
        const double beta_prev = beta;
        SCIRS_HIP_CHECK(hipMemcpyAsync(&beta, hdev, sizeof(double), hipMemcpyDeviceToHost, st.s));
        SCIRS_HIP_CHECK(hipStreamSynchronize(st.s));
        beta = std::sqrt(beta);
        if (!(beta > 0)) break;
        if (!(beta < 0.5 * beta_prev)) { *stalled = true; break; }
    }
    *iters = total;
    return SCIRS_OK;
}

int lowp_mode_of(int mode)
{
    switch (mode) {
    case SCIRS_MODE_BF16: return SCIRS_LOWP_BF16;
    case SCIRS_MODE_FP16: return SCIRS_LOWP_FP16;
    case SCIRS_MODE_BF16X3: return SCIRS_LOWP_BF16X3;
    case SCIRS_MODE_FP32: return SCIRS_LOWP_FP32;
    case SCIRS_MODE_BF16X6: return SCIRS_LOWP_BF16X6;
    default: return -1;
    }
}

} // namespace

namespace scirs {

int gesv_ir_device(int n, int nrhs, const double *dA, long lda, const double *dB, long ldb, double *dX, long ldx,
                   int mode, int max_iter, int max_gmres, double tol, hipStream_t s, scirs_solve_info *info)
{
    Context &c = ctx();
    int rc;
    if (mode == SCIRS_MODE_AUTO) {
        const char *m = getenv("SCIRS_GPU_MIXED_MODE");
        mode = m && *m ? atoi(m) : SCIRS_MODE_BF16X6;
    }
    const int lowp = lowp_mode_of(mode);
    if (lowp < 0) {
        scirs_set_error("dgesv_ir: unknown mode %d", mode);
        return SCIRS_ERR_ARG;
    }
    if (max_iter <= 0) max_iter = 30;
    if (max_gmres <= 0) max_gmres = 300;
    if (!(tol > 0)) tol = n * std::ldexp(1.0, -53);
    const int restart = std::min(n, std::max(1, atoi(getenv("SCIRS_GPU_GMRES_RESTART") ? getenv("SCIRS_GPU_GMRES_RESTART") : "50")));

    // scratch: A32 (n*n f32) | R (n*nrhs) | d32 (n) | norms (256) | rowsum (n) | V,Z ((restart+1)*n each) | w | h
    // the FP32 copy's rows start on 256-byte boundaries (16-byte loads of the tournament panel)
    const long ld32 = ((long)n + 63) / 64 * 64;
    const size_t a32 = (size_t)n * ld32 * sizeof(float);
    const size_t dinv_elems = (size_t)((n + 63) / 64) * 64 * 64;
    const size_t other = ((size_t)n * nrhs + (size_t)n + 256 + (size_t)n + 2 * (size_t)(restart + 1) * n + n + restart + 8) *
                         sizeof(double) + (size_t)n * sizeof(float) + (size_t)n * sizeof(int64_t) +
                         4 * dinv_elems * sizeof(float) + 4096;
    if ((rc = c.ensure(&c.scratch_ir, &c.scratch_ir_bytes, a32))) return rc;
    if ((rc = c.ensure(&c.scratch_ir2, &c.scratch_ir2_bytes, other))) return rc;
    IrState st;
    st.n = n; st.nrhs = nrhs; st.A = dA; st.lda = lda; st.B = dB; st.ldb = ldb; st.X = dX; st.ldx = ldx; st.s = s;
    st.LU = (float *)c.scratch_ir;
    st.ld32 = ld32;
    char *p = (char *)c.scratch_ir2;
    auto take = [&](size_t bytes) { char *r = p; p += (bytes + 255) & ~(size_t)255; return r; };
    st.R = (double *)take((size_t)n * nrhs * sizeof(double));
    st.norms = (double *)take(256 * sizeof(double));
    st.rowsum = (double *)take((size_t)n * sizeof(double));
    st.d32 = (float *)take((size_t)n * sizeof(float));
    st.perm = (int64_t *)take((size_t)n * sizeof(int64_t));
    double *V = (double *)take((size_t)(restart + 1) * n * sizeof(double));
    double *Z = (double *)take((size_t)(restart + 1) * n * sizeof(double));
    double *w = (double *)take((size_t)n * sizeof(double));
    double *hdev = (double *)take((size_t)(restart + 8) * sizeof(double));
    float *dinv_l = (float *)take(dinv_elems * sizeof(float));
    float *dinv_u = (float *)take(dinv_elems * sizeof(float));
    float *dmt_l = (float *)take(dinv_elems * sizeof(float));
    float *dmt_u = (float *)take(dinv_elems * sizeof(float));
    const bool tslu = c.work.tslu != 0;
    st.kb = tslu ? c.kb_f32 : c.lu_kb<float>(n);

    scirs::CallEvents ev;
    if (int erc = ev.create(3)) return erc;
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1], e2 = ev.e[2];
    SCIRS_HIP_CHECK(hipEventRecord(e0, s));
    // 1. FP32 copy of A and ||A||_inf; ||B_j||_inf
    if ((rc = scirs_convert_rowsum(n, n, dA, lda, st.LU, ld32, st.rowsum, s))) return rc;
    if ((rc = scirs_rowsum_max(n, st.rowsum, st.norms + 192, s))) return rc;
    if ((rc = scirs_colmax_d(n, nrhs, dB, ldb, st.norms + 128, s))) return rc;
    // 2. low-precision LU
    if ((rc = c.reset_flags(s))) return rc;
    if (lowp == SCIRS_LOWP_FP16) {
        if ((rc = c.ensure(&c.colexp_mem, &c.colexp_bytes, 2 * (size_t)n * sizeof(int)))) return rc;
        c.work.colexp = (int *)c.colexp_mem;
        c.work.colexp_half = n;
    }
    const int saved_mode = c.work.lowp_mode;
    c.work.lowp_mode = lowp;
    c.work.tslu_active = tslu;
    rc = scirs_getrf_device<float>(n, n, st.LU, ld32, st.kb, c.work, s);
    c.work.tslu_active = 0;
    c.work.lowp_mode = saved_mode;
    if (rc) return rc;
    if ((rc = scirs_perm_from_pairs(n, n, st.kb, c.work, st.perm, s))) return rc;
    // the FP32 solves apply inverted 64 x 64 diagonal blocks instead of the
    // dependent substitution inside each block (the factors only precondition)
    if (c.trsv_dinv) {
        if ((rc = scirs_tri_block_inverse(n, st.LU, ld32, 1, 1, dinv_l, s))) return rc;
        if ((rc = scirs_tri_block_inverse(n, st.LU, ld32, 0, 0, dinv_u, s))) return rc;
        if ((rc = scirs_tri_block_coupling(n, st.LU, ld32, 1, dinv_l, dmt_l, s))) return rc;
        if ((rc = scirs_tri_block_coupling(n, st.LU, ld32, 0, dinv_u, dmt_u, s))) return rc;
        st.dinv_l = dinv_l;
        st.dinv_u = dinv_u;
        st.dmt_l = dmt_l;
        st.dmt_u = dmt_u;
    }
    SCIRS_HIP_CHECK(hipEventRecord(e1, s));
    int lu_info = 0;
    double pv = 0;
    if ((rc = c.read_flags(s, &lu_info, &pv))) return rc;
    {
        double h[65];
        SCIRS_HIP_CHECK(hipMemcpyAsync(h, st.norms + 128, 65 * sizeof(double), hipMemcpyDeviceToHost, s));
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
        st.normB.assign(h, h + nrhs);
        st.normA = h[64];
    }
    int path = SCIRS_PATH_MIXED_IR;
    int ir_iters = 0, gm_iters = 0;
    double worst = INFINITY;
    bool converged = false;
    if (lu_info == 0) {
        // 3. refinement from X = 0
        hipLaunchKernelGGL(zero_ld_kernel, dim3(nb((long)n * nrhs)), dim3(256), 0, s, n, nrhs, dX, ldx);
        std::vector<double> eta, rn, prev(nrhs, INFINITY);
        std::vector<int> use_gmres(nrhs, 0), gm_col(nrhs, 0);   // GMRES budget per right-hand side
        for (int it = 0; it <= max_iter; ++it) {
            if ((rc = backward_errors(st, eta, rn))) return rc;
            worst = 0;
            for (double e : eta) worst = std::max(worst, e);
            if (worst <= tol) { converged = true; break; }
            bool budget_left = false;
            for (int j = 0; j < nrhs; ++j) budget_left = budget_left || (eta[j] > tol && gm_col[j] < max_gmres);
            if (it == max_iter || !budget_left) break;
            ++ir_iters;
            for (int j = 0; j < nrhs; ++j) {
                if (eta[j] <= tol) continue;
                // a step that does not halve eta marks the column for GMRES-IR
                if (it > 0 && !(eta[j] < 0.5 * prev[j])) use_gmres[j] = 1;
                if (use_gmres[j]) {
                    int used = 0;
                    path = SCIRS_PATH_MIXED_GMRES;
                    const double rtol = std::max(1e-12, std::min(0.5, tol * 16 / std::max(eta[j], 1e-300)));
                    if (gm_col[j] >= max_gmres) continue;
                    bool stalled = false;
                    if ((rc = gmres_correction(st, j, restart, max_gmres - gm_col[j], std::max(rtol, 1e-10), V, Z, w,
                                               hdev, &used, &stalled)))
                        return rc;
                    gm_iters += used;
                    gm_col[j] += used;
                    if (stalled) gm_col[j] = max_gmres;   // no further GMRES for this column
                } else {
                    if ((rc = ir_step(st, j, rn[j]))) return rc;
                }
            }
            prev = eta;
        }
    }
    if (!converged) {
        // 5. FP64 LU decides (singularity is judged the FP64 way)
        path = SCIRS_PATH_FP64_FALLBACK;
        double *dLU64 = nullptr;
        if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, (size_t)n * n * sizeof(double)))) return rc;
        dLU64 = (double *)c.scratch_a;
        if ((rc = scirs_d2d_copy(n, n, dA, lda, dLU64, n, s))) return rc;
        if ((rc = c.reset_flags(s))) return rc;
        const int kb64 = c.lu_kb<double>(n);
        if ((rc = scirs_getrf_device<double>(n, n, dLU64, n, kb64, c.work, s))) return rc;
        if ((rc = scirs_perm_from_pairs(n, n, kb64, c.work, st.perm, s))) return rc;
        if ((rc = scirs_permute_rows<double>(n, nrhs, st.perm, dB, ldb, dX, ldx, s))) return rc;
        if ((rc = scirs_getrs_device<double>(n, nrhs, dLU64, n, dX, ldx, c.work.info, s))) return rc;
        int st64 = 0;
        if ((rc = c.read_flags(s, &st64, &pv))) return rc;
        if (st64 != 0) {
            if (info) { info->pivot_value = pv; info->path = path; }
            return st64;
        }
        std::vector<double> eta, rn;
        if ((rc = backward_errors(st, eta, rn))) return rc;
        worst = 0;
        for (double e : eta) worst = std::max(worst, e);
    }
    SCIRS_HIP_CHECK(hipEventRecord(e2, s));
    SCIRS_HIP_CHECK(hipEventSynchronize(e2));
    if (info) {
        float ms1 = 0, ms2 = 0;
        (void)hipEventElapsedTime(&ms1, e0, e1);
        (void)hipEventElapsedTime(&ms2, e1, e2);
        info->time_factor_ms = ms1;
        info->time_solve_ms = ms2;
        info->path = path;
        info->ir_iters = ir_iters;
        info->gmres_iters = gm_iters;
        info->backward_error = worst;
    }
    return SCIRS_OK;
}

} // namespace scirs

using namespace scirs;

extern "C" int scirs_gpu_dgesv_ir_dev(int n, int nrhs, const double *dA, int64_t lda, const double *dB, int64_t ldb,
                                      double *dX, int64_t ldx, int mode, int max_iter, int max_gmres, double tol,
                                      void *stream, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_MIXED_IR);
    if (n <= 0 || nrhs <= 0 || nrhs > 64 || !dA || !dB || !dX || lda < n || ldb < nrhs || ldx < nrhs) {
        scirs_set_error("dgesv_ir_dev: invalid arguments (n=%d nrhs=%d; nrhs <= 64)", n, nrhs);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    int rc = ctx().init(0);
    if (rc) return finish(info, rc, t0);
    hipStream_t s = (hipStream_t)stream;   // NULL: the default stream, ordered after the caller's work on it
    rc = gesv_ir_device(n, nrhs, dA, lda, dB, ldb, dX, ldx, mode, max_iter, max_gmres, tol, s, info);
    return finish(info, rc, t0);
}

extern "C" int scirs_gpu_dgesv_ir(int n, int nrhs, const double *a, int64_t lda, const double *b, int64_t ldb,
                                  double *x, int64_t ldx, int mode, int max_iter, int max_gmres, double tol,
                                  scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_MIXED_IR);
    if (n <= 0 || nrhs <= 0 || nrhs > 64 || !a || !b || !x || lda < n || ldb < nrhs || ldx < nrhs) {
        scirs_set_error("dgesv_ir: invalid arguments (n=%d nrhs=%d; nrhs <= 64)", n, nrhs);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    Context &c = ctx();
    int rc = c.init(0);
    if (rc) return finish(info, rc, t0);
    hipStream_t s = c.stream;
    // keep the FP64 copy of A and B/X in scratch_b (scratch_a is used by the fallback)
    const size_t abytes = (size_t)n * n * sizeof(double), bbytes = (size_t)n * nrhs * sizeof(double);
    if ((rc = c.ensure(&c.scratch_b, &c.scratch_b_bytes, abytes + 2 * bbytes))) return finish(info, rc, t0);
    double *dA = (double *)c.scratch_b, *dB = dA + (size_t)n * n, *dX = dB + (size_t)n * nrhs;
    SCIRS_HIP_CHECK(hipMemcpy2DAsync(dA, n * sizeof(double), a, lda * sizeof(double), n * sizeof(double), n,
                                     hipMemcpyHostToDevice, s));
    SCIRS_HIP_CHECK(hipMemcpy2DAsync(dB, nrhs * sizeof(double), b, ldb * sizeof(double), nrhs * sizeof(double), n,
                                     hipMemcpyHostToDevice, s));
    rc = gesv_ir_device(n, nrhs, dA, n, dB, nrhs, dX, nrhs, mode, max_iter, max_gmres, tol, s, info);
    if (rc == 0) {
        SCIRS_HIP_CHECK(hipMemcpy2DAsync(x, ldx * sizeof(double), dX, nrhs * sizeof(double), nrhs * sizeof(double), n,
                                         hipMemcpyDeviceToHost, s));
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    }
    return finish(info, rc, t0);
}
