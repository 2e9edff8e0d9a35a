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
#include "prof.hpp"

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
    scirs::ProfScope ps(SCIRS_PROF_TRSV, (double)n * n * nrhs,
                        (0.5 * n * n + 2.0 * n * nrhs) * sizeof(T), s);
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

// ---------------------------------------------------------------------------
// Big-block triangular solve: 256-row diagonal blocks, two launches each.
//
//   big_diag:   one workgroup solves the 256 x 256 diagonal block in place,
//               in 64-row sub-blocks: one wave per right-hand side does the
//               substitution of a 64 x 64 sub-triangle staged in LDS (row i
//               broadcast from lane i), then all waves subtract that
//               sub-block's contribution from the block's later rows.
//   big_update: the rows outside the block take the block's contribution,
//               one wave per row, lanes across the 256 columns (coalesced).
// Same operations as the substitution loops of scirs2-linalg/src/solve.rs:159-199,
// with the k-sum split into blocks.
constexpr int BB = 256;     // big block
constexpr int SB = 64;      // sub-block
constexpr int QB = 4;       // right-hand sides per pass

template <typename T, bool LOWER, bool UNIT>
__global__ __launch_bounds__(NT) void big_diag_kernel(const T *A, long lda, T *X, long ldx, int nrhs, int b0, int bs,
                                                      const int *info)
{
    if (info && *info != 0) return;
    __shared__ T xs[BB][QB + 1];
    __shared__ T Ls[SB][SB + 1];          // the 64 x 64 sub-triangle
    __shared__ T Lr[BB - SB][SB + 1];     // rows of the block that the sub-block updates
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nsub = (bs + SB - 1) / SB;
    for (int rb = 0; rb < nrhs; rb += QB) {
        const int nr = min(QB, nrhs - rb);
        for (int idx = t; idx < bs * QB; idx += NT) {
            const int i = idx / QB, q = idx % QB;
            xs[i][q] = q < nr ? X[(long)(b0 + i) * ldx + rb + q] : (T)0;
        }
        for (int k = 0; k < nsub; ++k) {
            const int sub = LOWER ? k : nsub - 1 - k;
            const int s0 = sub * SB, ss = min(SB, bs - s0);
            const int rlo = LOWER ? s0 + ss : 0, rhi = LOWER ? bs : s0;
            __syncthreads();
            // stage the sub-triangle and the block rows it updates (coalesced rows)
            g2s_block<T>(&Ls[0][0], SB + 1, A + (long)(b0 + s0) * lda + b0 + s0, lda, SB, SB, ss, ss);
            if (rhi > rlo)
                g2s_block<T>(&Lr[0][0], SB + 1, A + (long)(b0 + rlo) * lda + b0 + s0, lda, rhi - rlo, SB, rhi - rlo, ss);
            __syncthreads();
            if (wave < nr) {
                const int q = wave;
                T b = lane < ss ? xs[s0 + lane][q] : (T)0;
                if (LOWER) {
                    for (int i = 0; i < ss; ++i) {
                        T xi = __shfl(b, i, 64);
                        if (!UNIT) xi = xi / Ls[i][i];
                        if (lane == i) b = xi;
                        if (lane > i && lane < ss) b -= Ls[lane][i] * xi;
                    }
                } else {
                    for (int i = ss - 1; i >= 0; --i) {
                        T xi = __shfl(b, i, 64);
                        if (!UNIT) xi = xi / Ls[i][i];
                        if (lane == i) b = xi;
                        if (lane < i) b -= Ls[lane][i] * xi;
                    }
                }
                if (lane < ss) xs[s0 + lane][q] = b;
            }
            __syncthreads();
            // later rows of the block: one thread per (row, right-hand side)
            for (int w = t; w < (rhi - rlo) * QB; w += NT) {
                const int r = rlo + w / QB, q = w % QB;
                if (q >= nr) continue;
                T acc = (T)0;
                for (int c = 0; c < ss; ++c) acc += Lr[r - rlo][c] * xs[s0 + c][q];
                xs[r][q] -= acc;
            }
        }
        __syncthreads();
        for (int idx = t; idx < bs * QB; idx += NT) {
            const int i = idx / QB, q = idx % QB;
            if (q < nr) X[(long)(b0 + i) * ldx + rb + q] = xs[i][q];
        }
        __syncthreads();
    }
}

// X[r, :] -= A[r, b0:b0+bs] X[b0:b0+bs, :] for r in [r0, r1).  Each wave takes
// four rows at once (their loads issue together), lanes across the columns.
template <typename T>
__global__ __launch_bounds__(NT) void big_update_kernel(const T *A, long lda, T *X, long ldx, int nrhs, int b0, int bs,
                                                        int r0, int r1, const int *info)
{
    if (info && *info != 0) return;
    __shared__ T xs[BB][QB + 1];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int RW = 4;                  // rows per wave
    constexpr int RPB = RW * (NT / 64);    // rows per workgroup
    const int rbase = r0 + blockIdx.x * RPB + wave * RW;
    for (int rb = 0; rb < nrhs; rb += QB) {
        const int nr = min(QB, nrhs - rb);
        __syncthreads();
        for (int idx = t; idx < bs * QB; idx += NT) {
            const int i = idx / QB, q = idx % QB;
            xs[i][q] = q < nr ? X[(long)(b0 + i) * ldx + rb + q] : (T)0;
        }
        __syncthreads();
        T acc[RW][QB];
#pragma unroll
        for (int a = 0; a < RW; ++a)
#pragma unroll
            for (int q = 0; q < QB; ++q) acc[a][q] = (T)0;
        for (int c = lane; c < bs; c += 64) {
            T l[RW];
#pragma unroll
            for (int a = 0; a < RW; ++a) l[a] = (rbase + a < r1) ? A[(long)(rbase + a) * lda + b0 + c] : (T)0;
#pragma unroll
            for (int a = 0; a < RW; ++a)
#pragma unroll
                for (int q = 0; q < QB; ++q) acc[a][q] += l[a] * xs[c][q];
        }
#pragma unroll
        for (int a = 0; a < RW; ++a)
#pragma unroll
            for (int q = 0; q < QB; ++q) acc[a][q] = wave_sum(acc[a][q]);
        if (lane < RW * QB) {
            const int a = lane / QB, q = lane % QB;
            T v = (T)0;
#pragma unroll
            for (int aa = 0; aa < RW; ++aa)
#pragma unroll
                for (int qq = 0; qq < QB; ++qq)
                    if (aa == a && qq == q) v = acc[aa][qq];
            if (rbase + a < r1 && q < nr) X[(long)(rbase + a) * ldx + rb + q] -= v;
        }
    }
}

} // namespace

template <typename T>
int scirs_trsv_big(int n, int nrhs, const T *A, long lda, T *X, long ldx, int lower, int unit, const int *info,
                   hipStream_t s)
{
    if (n <= 0 || nrhs <= 0) return SCIRS_OK;
    scirs::ProfScope prof(SCIRS_PROF_TRSV, (double)n * n * nrhs, (0.5 * n * n + 2.0 * n * nrhs) * sizeof(T), s);
    const int nblk = (n + BB - 1) / BB;
    for (int k = 0; k < nblk; ++k) {
        const int blk = lower ? k : nblk - 1 - k;
        const int b0 = blk * BB, bs = min(BB, n - b0);
        if (lower) {
            if (unit) hipLaunchKernelGGL((big_diag_kernel<T, true, true>), dim3(1), dim3(NT), 0, s, A, lda, X, ldx, nrhs, b0, bs, info);
            else hipLaunchKernelGGL((big_diag_kernel<T, true, false>), dim3(1), dim3(NT), 0, s, A, lda, X, ldx, nrhs, b0, bs, info);
        } else {
            if (unit) hipLaunchKernelGGL((big_diag_kernel<T, false, true>), dim3(1), dim3(NT), 0, s, A, lda, X, ldx, nrhs, b0, bs, info);
            else hipLaunchKernelGGL((big_diag_kernel<T, false, false>), dim3(1), dim3(NT), 0, s, A, lda, X, ldx, nrhs, b0, bs, info);
        }
        const int r0 = lower ? b0 + bs : 0, r1 = lower ? n : b0;
        if (r1 > r0)
            hipLaunchKernelGGL(big_update_kernel<T>, dim3((r1 - r0 + 15) / 16), dim3(NT), 0, s, A, lda, X, ldx, nrhs,
                               b0, bs, r0, r1, info);
    }
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

template int scirs_trsv_big<float>(int, int, const float *, long, float *, long, int, int, const int *, hipStream_t);
template int scirs_trsv_big<double>(int, int, const double *, long, double *, long, int, int, const int *, hipStream_t);

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
