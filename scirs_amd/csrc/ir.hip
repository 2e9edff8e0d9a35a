// Mixed-precision solve: factor once in low precision, refine in FP64.
//
// Counterpart of scirs2-linalg/src/mixed_precision/adaptive.rs:282-378
// (`iterative_refinement_solve`), which re-factors A in the working precision
// on every step and stops on an absolute residual.  Here:
//   1. A32 = fp32(A) fused with the row sums of |A| (||A||_inf);
//   2. LU of A32: FP32 panel with tournament pivoting, trailing update in the
//      requested arithmetic (BF16 / FP16-scaled / split-BF16 / split-FP16 /
//      FP32 matrix cores, FP32 accumulation);
//   3. refinement  R = B - A X (FP64, GPU), D = U^-1 L^-1 P R (FP32),
//      X += D, stopping when the normwise backward error
//      max_j ||R_j||_inf / (||A||_inf ||X_j||_inf + ||B_j||_inf) <= tol;
//   4. when a step fails to halve the backward error, restarted GMRES (FP64
//      Arnoldi with classical Gram-Schmidt applied twice, right-preconditioned
//      by the low-precision factors upcast to FP64, up to 8 right-hand sides in
//      lockstep) solves the correction equation A D = R
//      (Carson & Higham, SISC 40(2), 2018, "GMRES-IR"); the Givens least-squares
//      update is the one of scirs2-linalg/src/solvers/iterative.rs:604-664;
//   5. when GMRES-IR also stalls, or the low-precision LU breaks down, FP64 LU
//      solves the system (the breakdown is then decided the FP64 way).
// FP64 is used on the GPU only for the residual, the update of X and the
// stopping test (plus the Arnoldi vectors and the preconditioner of GMRES).
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

// ---- block GMRES: nc right-hand sides side by side.  A block of vectors is
// n x nc, row-major (entry (q, c) at q * nc + c), so the triangular solves and
// the product with A read the matrix once for all nc columns.
constexpr int GM_NC = 8;   // columns per block GMRES pass

// W = A Z for nc <= GM_NC columns; one wave per 8 rows of A
__global__ __launch_bounds__(NT) void gemv_block_kernel(int n, int nc, const double *A, long lda, const double *Z,
                                                        double *W)
{
    const int lane = threadIdx.x & 63;
    const int r0 = ((blockIdx.x * NT + threadIdx.x) >> 6) * 8;
    if (r0 >= n) return;
    const int nr = min(8, n - r0);
    double acc[8][GM_NC];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < GM_NC; ++c) acc[a][c] = 0.0;
    for (int k = lane; k < n; k += 64) {
        double z[GM_NC];
#pragma unroll
        for (int c = 0; c < GM_NC; ++c) z[c] = c < nc ? Z[(long)k * nc + c] : 0.0;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const double av = a < nr ? A[(long)(r0 + a) * lda + k] : 0.0;
#pragma unroll
            for (int c = 0; c < GM_NC; ++c) acc[a][c] = fma(av, z[c], acc[a][c]);
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < GM_NC; ++c) acc[a][c] = wave_sum(acc[a][c]);
    if (lane == 0)
        for (int a = 0; a < nr; ++a)
            for (int c = 0; c < nc; ++c) W[(long)(r0 + a) * nc + c] = acc[a][c];
}

// h[i * nc + c] = <V_i(:, c), W(:, c)> for i < k (V_i: block i of n x nc)
__global__ __launch_bounds__(NT) void multidot_block_kernel(int n, int nc, const double *V, const double *W, double *h)
{
    const int i = blockIdx.x;
    const double *v = V + (long)i * n * nc;
    double s[GM_NC];
#pragma unroll
    for (int c = 0; c < GM_NC; ++c) s[c] = 0.0;
    for (int q = threadIdx.x; q < n; q += NT)
#pragma unroll
        for (int c = 0; c < GM_NC; ++c)
            if (c < nc) s[c] = fma(v[(long)q * nc + c], W[(long)q * nc + c], s[c]);
    __shared__ double red[NT / 64][GM_NC];
#pragma unroll
    for (int c = 0; c < GM_NC; ++c) {
        const double w = wave_sum(s[c]);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = w;
    }
    __syncthreads();
    if (threadIdx.x < nc) {
        double r = 0;
        for (int q = 0; q < NT / 64; ++q) r += red[q][threadIdx.x];
        h[(long)i * nc + threadIdx.x] = r;
    }
}

// W(:, c) -= sum_{i < k} h[i * nc + c] V_i(:, c)
__global__ void mgs_block_kernel(int n, int nc, int k, const double *V, const double *h, double *W)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * nc) return;
    const int c = (int)(e % nc);
    double s = W[e];
    for (int i = 0; i < k; ++i) s -= h[(long)i * nc + c] * V[(long)i * n * nc + e];
    W[e] = s;
}

// dst(:, c) = src(:, c) / sqrt(ss[c])  (0 where ss[c] is 0: a breakdown column stays 0)
__global__ void normalize_block_kernel(int n, int nc, const double *src, const double *ss, double *dst)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * nc) return;
    const double v = ss[e % nc];
    dst[e] = v > 0 ? src[e] / sqrt(v) : 0.0;
}

// V0(:, c) = R(:, cols[c])
__global__ void gather_cols_kernel(int n, int nc, const double *R, long ldr, const int *cols, double *V0)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * nc) return;
    const long q = e / nc;
    V0[e] = R[q * ldr + cols[e % nc]];
}

// X(:, cols[c]) += sum_{j < kmax} Y[j * nc + c] Z_j(:, c)
__global__ void combine_block_kernel(int n, int nc, int kmax, const double *Z, const double *Y, double *X, long ldx,
                                     const int *cols)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * nc) return;
    const int c = (int)(e % nc);
    double s = 0.0;
    for (int j = 0; j < kmax; ++j) s += Y[(long)j * nc + c] * Z[(long)j * n * nc + e];
    X[(e / nc) * ldx + cols[c]] += s;
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

// eta_j for every column; returns the worst, fills per-column values
int backward_errors(IrState &st, std::vector<double> &eta, std::vector<double> &rnorm, bool x_zero = false)
{
    int rc;
    if (x_zero) {
        // X = 0: R = B, ||R_j|| = ||B_j|| and eta_j = 1 (0 for a zero column),
        // with no pass over A
        if ((rc = scirs_d2d_copy(st.n, st.nrhs, st.B, st.ldb, st.R, st.nrhs, st.s))) return rc;
        eta.assign(st.nrhs, 0.0);
        rnorm.assign(st.nrhs, 0.0);
        for (int j = 0; j < st.nrhs; ++j) {
            rnorm[j] = st.normB[j];
            eta[j] = st.normB[j] > 0 ? 1.0 : 0.0;
        }
        return SCIRS_OK;
    }
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

// Restarted GMRES on A D = R(:, cols) for nc <= GM_NC columns side by side,
// right-preconditioned by the low-precision factors applied in FP64:
// z = U^-1 L^-1 P v with the factors upcast once to FP64 (LU64), so the
// preconditioner adds no rounding of its own above 2^-53 (Amestoy, Buttari,
// Higham, L'Excellent, Mary, Vieuble, SIMAX 2024: the precision of the
// preconditioner's application bounds what GMRES-IR can resolve).  Each
// column keeps its own Hessenberg matrix and Givens rotations (the small
// least-squares update of scirs2-linalg/src/solvers/iterative.rs:604-664); the
// device work of one Arnoldi step — preconditioner, product with A, classical
// Gram-Schmidt twice, normalisation — is queued for all columns at once and
// followed by one host read (the new Hessenberg columns).  A column stops
// when its Arnoldi residual estimate reaches rtol times its starting residual,
// when its budget is spent, or when the contraction of its true residual over
// the last restart cycle, extrapolated, needs more cycles than its remaining
// budget holds (stalled[c] is then set: the caller moves on at once instead of
// spending the budget on a preconditioner too weak for this system).
struct GmresWork {
    double *V, *Z, *W, *h, *Y;   // (restart+1) blocks of n x nc each for V and Z; W n x nc
    int *cols;                   // device copy of the column indices
    double *LU64;                // FP64 copy of the factors
};

int precond_block(IrState &st, const GmresWork &gw, int nc, const double *Vk, double *Zk)
{
    int rc;
    if ((rc = scirs_permute_rows<double>(st.n, nc, st.perm, Vk, nc, Zk, nc, st.s))) return rc;
    return scirs::lu_substitute<double>(st.n, nc, gw.LU64, st.n, Zk, nc, st.s);
}

int gmres_block(IrState &st, const std::vector<int> &cols, int restart, const std::vector<int> &budget,
                const std::vector<double> &rtol, const GmresWork &gw, std::vector<int> &used, std::vector<char> &stalled)
{
    const int n = st.n, nc = (int)cols.size();
    const long blk = (long)n * nc;
    const unsigned nbe = nb(blk);
    int rc;
    used.assign(nc, 0);
    stalled.assign(nc, 0);
    std::vector<char> active(nc, 1);
    SCIRS_HIP_CHECK(hipMemcpyAsync(gw.cols, cols.data(), nc * sizeof(int), hipMemcpyHostToDevice, st.s));
    std::vector<double> H((size_t)nc * (restart + 1) * restart), cs((size_t)nc * restart), sn((size_t)nc * restart),
        g((size_t)nc * (restart + 1)), hbuf((size_t)(2 * (restart + 1) + 1) * nc), beta(nc), beta_prev(nc, INFINITY),
        target(nc, 0.0), Y((size_t)restart * nc);
    std::vector<int> kc(nc);
    bool first = true;
    while (true) {
        // V_0 = R(:, cols) / ||R(:, cols)||, one host read for the norms
        hipLaunchKernelGGL(gather_cols_kernel, dim3(nbe), dim3(256), 0, st.s, n, nc, st.R, (long)st.nrhs, gw.cols, gw.V);
        hipLaunchKernelGGL(multidot_block_kernel, dim3(1), dim3(NT), 0, st.s, n, nc, gw.V, gw.V, gw.h);
        SCIRS_HIP_CHECK(hipMemcpyAsync(hbuf.data(), gw.h, nc * sizeof(double), hipMemcpyDeviceToHost, st.s));
        SCIRS_HIP_CHECK(hipStreamSynchronize(st.s));
        bool any = false;
        for (int c = 0; c < nc; ++c) {
            beta[c] = std::sqrt(hbuf[c]);
            if (!active[c]) continue;
            if (!(beta[c] > 0)) { active[c] = 0; continue; }
            if (first) target[c] = rtol[c] * beta[c];
            else if (beta[c] <= target[c]) { active[c] = 0; continue; }
            else {
                // the last cycle's contraction, extrapolated: give up at once when
                // the cycles it predicts do not fit the remaining budget
                const double ratio = beta[c] / beta_prev[c];
                const double need = ratio < 1.0 ? std::log(target[c] / beta[c]) / std::log(ratio) : INFINITY;
                if (!(need * restart <= budget[c] - used[c])) { stalled[c] = 1; active[c] = 0; continue; }
            }
            if (used[c] >= budget[c]) { active[c] = 0; continue; }
            beta_prev[c] = beta[c];
            any = true;
        }
        first = false;
        if (!any) break;
        hipLaunchKernelGGL(normalize_block_kernel, dim3(nbe), dim3(256), 0, st.s, n, nc, gw.V, gw.h, gw.V);
        std::vector<char> run(active);   // columns still iterating in this cycle
        for (int c = 0; c < nc; ++c) {
            std::fill(g.begin() + (size_t)c * (restart + 1), g.begin() + (size_t)(c + 1) * (restart + 1), 0.0);
            g[(size_t)c * (restart + 1)] = beta[c];
            kc[c] = 0;
        }
        int k = 0;
        for (; k < restart; ++k) {
            bool anyrun = false;
            for (int c = 0; c < nc; ++c) anyrun = anyrun || run[c];
            if (!anyrun) break;
            const double *vk = gw.V + (long)k * blk;
            double *vk1 = gw.V + (long)(k + 1) * blk, *zk = gw.Z + (long)k * blk;
            if ((rc = precond_block(st, gw, nc, vk, zk))) return rc;
            hipLaunchKernelGGL(gemv_block_kernel, dim3(nb((long)((n + 7) / 8) * 64, NT)), dim3(NT), 0, st.s, n, nc,
                               st.A, st.lda, zk, gw.W);
            // classical Gram-Schmidt, twice; h rows 0..k of pass 1 and 2, then ||w||^2
            double *h1 = gw.h, *h2 = gw.h + (long)(restart + 1) * nc, *hw = gw.h + 2L * (restart + 1) * nc;
            hipLaunchKernelGGL(multidot_block_kernel, dim3(k + 1), dim3(NT), 0, st.s, n, nc, gw.V, gw.W, h1);
            hipLaunchKernelGGL(mgs_block_kernel, dim3(nbe), dim3(256), 0, st.s, n, nc, k + 1, gw.V, h1, gw.W);
            hipLaunchKernelGGL(multidot_block_kernel, dim3(k + 1), dim3(NT), 0, st.s, n, nc, gw.V, gw.W, h2);
            hipLaunchKernelGGL(mgs_block_kernel, dim3(nbe), dim3(256), 0, st.s, n, nc, k + 1, gw.V, h2, gw.W);
            hipLaunchKernelGGL(multidot_block_kernel, dim3(1), dim3(NT), 0, st.s, n, nc, gw.W, gw.W, hw);
            hipLaunchKernelGGL(normalize_block_kernel, dim3(nbe), dim3(256), 0, st.s, n, nc, gw.W, hw, vk1);
            SCIRS_HIP_CHECK(hipGetLastError());
            SCIRS_HIP_CHECK(hipMemcpyAsync(hbuf.data(), gw.h, (size_t)(2 * (restart + 1) + 1) * nc * sizeof(double),
                                           hipMemcpyDeviceToHost, st.s));
            SCIRS_HIP_CHECK(hipStreamSynchronize(st.s));
            for (int c = 0; c < nc; ++c) {
                if (!run[c]) continue;
                double *Hc = &H[(size_t)c * (restart + 1) * restart];
                double *csc = &cs[(size_t)c * restart], *snc = &sn[(size_t)c * restart];
                double *gc = &g[(size_t)c * (restart + 1)];
                std::vector<double> hcol(k + 2);
                for (int i = 0; i <= k; ++i)
                    hcol[i] = hbuf[(size_t)i * nc + c] + hbuf[(size_t)(restart + 1) * nc + (size_t)i * nc + c];
                const double hn = std::sqrt(std::max(0.0, hbuf[(size_t)2 * (restart + 1) * nc + c]));
                hcol[k + 1] = hn;
                for (int i = 0; i < k; ++i) {
                    const double a = hcol[i], b = hcol[i + 1];
                    hcol[i] = csc[i] * a + snc[i] * b;
                    hcol[i + 1] = -snc[i] * a + csc[i] * b;
                }
                const double a = hcol[k], b = hcol[k + 1];
                const double r = std::hypot(a, b);
                csc[k] = r > 0 ? a / r : 1.0;
                snc[k] = r > 0 ? b / r : 0.0;
                hcol[k] = r;
                gc[k + 1] = -snc[k] * gc[k];
                gc[k] = csc[k] * gc[k];
                for (int i = 0; i <= k; ++i) Hc[(size_t)i * restart + k] = hcol[i];
                kc[c] = k + 1;
                ++used[c];
                if (std::fabs(gc[k + 1]) <= target[c] || !(hn > 0) || used[c] >= budget[c] || !(r > 0))
                    run[c] = 0;
            }
        }
        // y_c by back substitution on each column's Hessenberg matrix, then X(:, cols) += Z y
        int kmax = 0;
        std::fill(Y.begin(), Y.end(), 0.0);
        for (int c = 0; c < nc; ++c) {
            if (!active[c] || kc[c] == 0) continue;
            const double *Hc = &H[(size_t)c * (restart + 1) * restart];
            const double *gc = &g[(size_t)c * (restart + 1)];
            int kk = kc[c];
            while (kk > 0 && !(Hc[(size_t)(kk - 1) * restart + kk - 1] != 0.0)) --kk;   // drop a breakdown step
            for (int i = kk - 1; i >= 0; --i) {
                double sacc = gc[i];
                for (int q = i + 1; q < kk; ++q) sacc -= Hc[(size_t)i * restart + q] * Y[(size_t)q * nc + c];
                Y[(size_t)i * nc + c] = sacc / Hc[(size_t)i * restart + i];
            }
            kmax = std::max(kmax, kk);
        }
        if (kmax > 0) {
            SCIRS_HIP_CHECK(hipMemcpyAsync(gw.Y, Y.data(), (size_t)kmax * nc * sizeof(double), hipMemcpyHostToDevice,
                                           st.s));
            hipLaunchKernelGGL(combine_block_kernel, dim3(nbe), dim3(256), 0, st.s, n, nc, kmax, gw.Z, gw.Y, st.X,
                               st.ldx, gw.cols);
            SCIRS_HIP_CHECK(hipGetLastError());
        }
        // true residuals of every column (one product with A), then the next cycle
        if ((rc = scirs_residual(n, st.nrhs, st.A, st.lda, st.X, st.ldx, st.B, st.ldb, st.R, st.nrhs, st.s))) return rc;
    }
    return SCIRS_OK;
}

int lowp_mode_of(int mode)
{
    switch (mode) {
    case SCIRS_MODE_BF16: return SCIRS_LOWP_BF16;
    case SCIRS_MODE_FP16: return SCIRS_LOWP_FP16;
    case SCIRS_MODE_FP16X3: return SCIRS_LOWP_FP16X3;
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
    const int gnc = std::min(nrhs, GM_NC);   // columns per block GMRES pass
    const size_t other = ((size_t)n * nrhs + (size_t)n + 256 + (size_t)n + 2 * (size_t)(restart + 1) * n * gnc +
                          (size_t)n * gnc + (size_t)(2 * (restart + 1) + 1) * gnc + (size_t)restart * gnc + 64) *
                             sizeof(double) +
                         (size_t)n * sizeof(float) + (size_t)n * sizeof(int64_t) + 8 * dinv_elems * sizeof(float) +
                         (size_t)gnc * sizeof(int) + 8192;
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
    GmresWork gw;
    gw.V = (double *)take((size_t)(restart + 1) * n * gnc * sizeof(double));
    gw.Z = (double *)take((size_t)(restart + 1) * n * gnc * sizeof(double));
    gw.W = (double *)take((size_t)n * gnc * sizeof(double));
    gw.h = (double *)take((size_t)(2 * (restart + 1) + 1) * gnc * sizeof(double));
    gw.Y = (double *)take((size_t)restart * gnc * sizeof(double));
    gw.cols = (int *)take((size_t)gnc * sizeof(int));
    gw.LU64 = nullptr;
    float *dinv_l = (float *)take(dinv_elems * sizeof(float));
    float *dinv_u = (float *)take(dinv_elems * sizeof(float));
    float *dmt_l = (float *)take(3 * dinv_elems * sizeof(float));   // couplings one to three blocks back
    float *dmt_u = (float *)take(3 * dinv_elems * sizeof(float));
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
    if (lowp == SCIRS_LOWP_FP16 || lowp == SCIRS_LOWP_FP16X3) {
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
        const int depth = c.work.trsv_pair ? 3 : 1;
        if ((rc = scirs_tri_block_coupling(n, st.LU, ld32, 1, dinv_l, dmt_l, depth, s))) return rc;
        if ((rc = scirs_tri_block_coupling(n, st.LU, ld32, 0, dinv_u, dmt_u, depth, s))) return rc;
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
            if ((rc = backward_errors(st, eta, rn, it == 0))) return rc;
            worst = 0;
            for (double e : eta) worst = std::max(worst, e);
            if (worst <= tol) { converged = true; break; }
            bool budget_left = false;
            for (int j = 0; j < nrhs; ++j) budget_left = budget_left || (eta[j] > tol && gm_col[j] < max_gmres);
            if (it == max_iter || !budget_left) break;
            ++ir_iters;
            std::vector<int> gcols, gbud;
            std::vector<double> grtol;
            for (int j = 0; j < nrhs; ++j) {
                if (eta[j] <= tol) continue;
                // a step that does not halve eta marks the column for GMRES-IR
                if (it > 0 && !(eta[j] < 0.5 * prev[j])) use_gmres[j] = 1;
                if (use_gmres[j]) {
                    if (gm_col[j] >= max_gmres) continue;
                    gcols.push_back(j);
                    gbud.push_back(max_gmres - gm_col[j]);
                    grtol.push_back(std::max(1e-10, std::min(0.5, tol * 16 / std::max(eta[j], 1e-300))));
                } else {
                    if ((rc = ir_step(st, j, rn[j]))) return rc;
                }
            }
            if (!gcols.empty()) {
                path = SCIRS_PATH_MIXED_GMRES;
                if (!gw.LU64) {
                    // the factors in FP64, once: the preconditioner is applied in FP64
                    if ((rc = c.ensure(&c.scratch_lu64, &c.scratch_lu64_bytes, (size_t)n * n * sizeof(double))))
                        return rc;
                    gw.LU64 = (double *)c.scratch_lu64;
                    if ((rc = scirs_s2d(n, n, st.LU, ld32, gw.LU64, n, s))) return rc;
                }
                for (size_t b0 = 0; b0 < gcols.size(); b0 += GM_NC) {
                    const size_t b1 = std::min(gcols.size(), b0 + GM_NC);
                    std::vector<int> cols(gcols.begin() + b0, gcols.begin() + b1), bud(gbud.begin() + b0,
                                                                                      gbud.begin() + b1);
                    std::vector<double> rt(grtol.begin() + b0, grtol.begin() + b1);
                    std::vector<int> used;
                    std::vector<char> stalled;
                    if ((rc = gmres_block(st, cols, restart, bud, rt, gw, used, stalled))) return rc;
                    for (size_t q = 0; q < cols.size(); ++q) {
                        gm_iters += used[q];
                        gm_col[cols[q]] += used[q];
                        if (stalled[q]) gm_col[cols[q]] = max_gmres;   // no further GMRES for this column
                    }
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
    if ((rc = scirs::copy_h2d(dA, n * sizeof(double), a, lda * sizeof(double), n * sizeof(double), n, s)))
        return finish(info, rc, t0);
    if ((rc = scirs::copy_h2d(dB, nrhs * sizeof(double), b, ldb * sizeof(double), nrhs * sizeof(double), n, s)))
        return finish(info, rc, t0);
    rc = gesv_ir_device(n, nrhs, dA, n, dB, nrhs, dX, nrhs, mode, max_iter, max_gmres, tol, s, info);
    if (rc == 0) {
        if ((rc = scirs::copy_d2h(x, ldx * sizeof(double), dX, nrhs * sizeof(double), nrhs * sizeof(double), n, s)))
            return finish(info, rc, t0);
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    }
    return finish(info, rc, t0);
}
