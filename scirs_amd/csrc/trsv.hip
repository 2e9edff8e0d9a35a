// Blocked triangular solves with several right-hand sides, in place on B
// (n x nrhs, row-major), for the solve phase after getrf / potrf:
//     LU:       L y = P b (unit lower), then U x = y (upper)
//     Cholesky: L y = b (lower), then L^T x = y (L^T read through L)
// The same arithmetic as scirs2-linalg/src/solve.rs:159-199 (forward and back
// substitution, true division by the diagonal) in 64-row blocks: a one-workgroup
// kernel solves the diagonal block (one wave per right-hand side, row i of the
// block is broadcast from lane i), then a many-workgroup kernel subtracts the
// block's contribution from every remaining row.
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int NBS = 64;
constexpr int NT = 256;

// element (i, j) of the triangular operand: L/U directly, or L^T through L
template <typename T, bool TRANS>
__device__ __forceinline__ T tri_at(const T *A, long lda, int i, int j)
{
    return TRANS ? A[(long)j * lda + i] : A[(long)i * lda + j];
}

// Solve the diagonal block [b0, b0+bs) for all right-hand sides.
template <typename T, bool LOWER, bool UNIT, bool TRANS>
__global__ __launch_bounds__(NT) void tri_diag_kernel(const T *A, long lda, T *B, long ldb, int nrhs,
                                                      int b0, int bs, const int *info)
{
    if (info && *info != 0) return;
    __shared__ T Ls[NBS][NBS + 1];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int idx = t; idx < bs * bs; idx += NT) {
        int i = idx / bs, j = idx % bs;
        Ls[i][j] = tri_at<T, TRANS>(A, lda, b0 + i, b0 + j);
    }
    __syncthreads();
    for (int rhs = wave; rhs < nrhs; rhs += NT / 64) {
        T b = lane < bs ? B[(long)(b0 + lane) * ldb + rhs] : (T)0;
        if (LOWER) {
            for (int i = 0; i < bs; ++i) {
                T xi = __shfl(b, i, 64);
                if (!UNIT) xi = xi / Ls[i][i];
                if (lane == i) b = xi;
                if (lane > i && lane < bs) b -= Ls[lane][i] * xi;
            }
        } else {
            for (int i = bs - 1; i >= 0; --i) {
                T xi = __shfl(b, i, 64);
                if (!UNIT) xi = xi / Ls[i][i];
                if (lane == i) b = xi;
                if (lane < i) b -= Ls[lane][i] * xi;
            }
        }
        if (lane < bs) B[(long)(b0 + lane) * ldb + rhs] = b;
    }
}

// B[rows r0..r1) -= Tri[rows, b0..b0+bs) * X[b0..b0+bs), X being rows of B.
template <typename T, bool TRANS>
__global__ __launch_bounds__(NT) void tri_update_kernel(const T *A, long lda, T *B, long ldb, int nrhs,
                                                        int b0, int bs, int r0, int r1, const int *info)
{
    if (info && *info != 0) return;
    __shared__ T Xs[NBS][17];
    const int t = threadIdx.x;
    const int rowbase = r0 + blockIdx.x * 64;
    for (int rb = 0; rb < nrhs; rb += 16) {
        const int nr = min(16, nrhs - rb);
        __syncthreads();
        for (int idx = t; idx < bs * nr; idx += NT) {
            int q = idx / nr, j = idx % nr;
            Xs[q][j] = B[(long)(b0 + q) * ldb + rb + j];
        }
        __syncthreads();
        // 4 threads per row, each taking every 4th right-hand side
        const int r = rowbase + (t >> 2);
        if (r < r1) {
            for (int j = (t & 3); j < nr; j += 4) {
                T s = (T)0;
                for (int q = 0; q < bs; ++q) s += tri_at<T, TRANS>(A, lda, r, b0 + q) * Xs[q][j];
                B[(long)r * ldb + rb + j] -= s;
            }
        }
    }
}

template <typename T, bool LOWER, bool UNIT, bool TRANS>
int tri_solve(int n, int nrhs, const T *A, long lda, T *B, long ldb, const int *info, hipStream_t s)
{
    const int nblk = (n + NBS - 1) / NBS;
    for (int k = 0; k < nblk; ++k) {
        const int blk = LOWER ? k : nblk - 1 - k;
        const int b0 = blk * NBS, bs = min(NBS, n - b0);
        hipLaunchKernelGGL((tri_diag_kernel<T, LOWER, UNIT, TRANS>), dim3(1), dim3(NT), 0, s,
                           A, lda, B, ldb, nrhs, b0, bs, info);
        int r0, r1;
        if (LOWER) { r0 = b0 + bs; r1 = n; } else { r0 = 0; r1 = b0; }
        if (r1 > r0)
            hipLaunchKernelGGL((tri_update_kernel<T, TRANS>), dim3((r1 - r0 + 63) / 64), dim3(NT), 0, s,
                               A, lda, B, ldb, nrhs, b0, bs, r0, r1, info);
    }
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

} // namespace

template <typename T>
int scirs_getrs_device(int n, int nrhs, const T *A, long lda, T *B, long ldb, const int *info,
                       hipStream_t stream)
{
    int rc = tri_solve<T, true, true, false>(n, nrhs, A, lda, B, ldb, info, stream);
    if (rc) return rc;
    return tri_solve<T, false, false, false>(n, nrhs, A, lda, B, ldb, info, stream);
}

template <typename T>
int scirs_potrs_device(int n, int nrhs, const T *L, long ldl, T *B, long ldb, const int *info,
                       hipStream_t stream)
{
    int rc = tri_solve<T, true, false, false>(n, nrhs, L, ldl, B, ldb, info, stream);
    if (rc) return rc;
    return tri_solve<T, false, false, true>(n, nrhs, L, ldl, B, ldb, info, stream);
}

template <typename T>
int scirs_trsv_device(int n, int nrhs, const T *A, long lda, T *B, long ldb, int lower, int unit,
                      const int *info, hipStream_t stream)
{
    if (lower)
        return unit ? tri_solve<T, true, true, false>(n, nrhs, A, lda, B, ldb, info, stream)
                    : tri_solve<T, true, false, false>(n, nrhs, A, lda, B, ldb, info, stream);
    return unit ? tri_solve<T, false, true, false>(n, nrhs, A, lda, B, ldb, info, stream)
                : tri_solve<T, false, false, false>(n, nrhs, A, lda, B, ldb, info, stream);
}

template int scirs_getrs_device<double>(int, int, const double *, long, double *, long, const int *, hipStream_t);
template int scirs_getrs_device<float>(int, int, const float *, long, float *, long, const int *, hipStream_t);
template int scirs_potrs_device<double>(int, int, const double *, long, double *, long, const int *, hipStream_t);
template int scirs_potrs_device<float>(int, int, const float *, long, float *, long, const int *, hipStream_t);
template int scirs_trsv_device<double>(int, int, const double *, long, double *, long, int, int, const int *, hipStream_t);
template int scirs_trsv_device<float>(int, int, const float *, long, float *, long, int, int, const int *, hipStream_t);
