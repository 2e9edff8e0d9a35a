// Blocked right-looking Cholesky (lower) for gfx950, row-major, in place.
// Only the lower triangle of A is read, as in scirs2-linalg/src/lapack.rs:632-679,
// and the strict upper triangle of the result is zeroed so the output is the
// reference's L.  Per block of kb columns at k0:
//   1. potf2 — one workgroup factors the kb x kb diagonal block in LDS; the
//      reference's test `a_ii - sum <= 0` (lapack.rs:661-670) sets
//      info = row + 1 and stops.
//   2. trsm  — L21 = A21 L11^-T, 64 rows per workgroup, L11 packed in LDS.
//   3. syrk  — A22 -= L21 L21^T on the matrix cores (lower tiles only).
#include "common.hpp"
#include "kernels.hpp"
#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace {
constexpr int NT = 256;

template <typename T>
__device__ __forceinline__ T readlane_t(T v, int k);
template <>
__device__ __forceinline__ float readlane_t<float>(float v, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
template <>
__device__ __forceinline__ double readlane_t<double>(double v, int k)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The kb x kb diagonal block (kb <= 128) in LDS, factored 16 columns at a time:
//   (a) wave 0 holds the 16 columns of every row in registers (lane = row mod 64,
//       two rows per lane) and factors them column by column with no barrier:
//       the diagonal by readlane, sqrt, true division of the rows below
//       (lapack.rs:661-676), the later columns of the 16 updated with the
//       column's entries broadcast by readlane;
//   (b) all waves apply the rank-16 update to the lower triangle right of and
//       below the 16 columns, in 4 x 4 register tiles.
// The test of the reference, a_jj - sum <= 0 (lapack.rs:661-670), sets
// info = row + 1 and stops.
template <typename T>
__global__ __launch_bounds__(NT) void potf2_kernel(T *A, long lda, int k0, int kb, int *info, double *info_val)
{
    if (*info != 0) return;
    constexpr int PB = 16;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // the lower triangle, packed by rows (row r at r (r + 1) / 2): half the LDS
    // of a square tile, so the kernel fits on a CU beside one workgroup of the
    // trailing update running on the other stream
    T *S = reinterpret_cast<T *>(smem_raw);
    auto pk = [](int r, int c) { return r * (r + 1) / 2 + c; };
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    T *D = A + (long)k0 * lda + k0;
    for (int i = wave; i < kb; i += NT / 64)
        for (int j = lane; j <= i; j += 64) S[pk(i, j)] = D[(long)i * lda + j];
    __shared__ int s_fail;
    if (t == 0) s_fail = 0;
    __syncthreads();
    for (int jb = 0; jb < kb; jb += PB) {
        const int pw = min(PB, kb - jb);
        if (wave == 0) {
            T a[2][PB];
#pragma unroll
            for (int sl = 0; sl < 2; ++sl)
#pragma unroll
                for (int i = 0; i < PB; ++i) {
                    const int row = lane + 64 * sl;
                    a[sl][i] = (row < kb && i < pw && jb + i <= row) ? S[pk(row, jb + i)] : (T)0;
                }
            int fail_row = -1;
            T fail_val = (T)0;
#pragma unroll
            for (int i = 0; i < PB; ++i) {
                if (i < pw && fail_row < 0) {
                    const int cj = jb + i;
                    const T dv = readlane_t<T>((cj >> 6) ? a[1][i] : a[0][i], cj & 63);
                    if (!(dv > (T)0)) {
                        fail_row = cj;
                        fail_val = dv;
                    } else {
                        const T d = sqrt(dv);
#pragma unroll
                        for (int sl = 0; sl < 2; ++sl) {
                            const int row = lane + 64 * sl;
                            if (row == cj) a[sl][i] = d;
                            else if (row > cj) a[sl][i] = a[sl][i] / d;
                        }
#pragma unroll
                        for (int i2 = i + 1; i2 < PB; ++i2) {
                            if (i2 < pw) {
                                const int c2 = jb + i2;
                                const T lc = readlane_t<T>((c2 >> 6) ? a[1][i] : a[0][i], c2 & 63);
#pragma unroll
                                for (int sl = 0; sl < 2; ++sl)
                                    if (lane + 64 * sl > cj) a[sl][i2] -= a[sl][i] * lc;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int sl = 0; sl < 2; ++sl)
#pragma unroll
                for (int i = 0; i < PB; ++i) {
                    const int row = lane + 64 * sl;
                    if (row < kb && i < pw && row >= jb + i) S[pk(row, jb + i)] = a[sl][i];
                }
            if (fail_row >= 0 && lane == 0) {
                s_fail = 1;
                *info = k0 + fail_row + 1;
                *info_val = (double)fail_val;
            }
        }
        __syncthreads();
        if (s_fail) return;
        // (b) S[r][c] -= sum_k S[r][jb+k] S[c][jb+k] for je <= c <= r < kb, 4 x 4 tiles
        const int je = jb + pw;
        const int nt = (kb - je + 3) / 4;
        const int ntiles = nt * (nt + 1) / 2;
        for (int tile = t; tile < ntiles; tile += NT) {
            int ti = (int)((sqrtf(8.0f * (float)tile + 1.0f) - 1.0f) * 0.5f);
            while (ti * (ti + 1) / 2 > tile) --ti;
            while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
            const int tj = tile - ti * (ti + 1) / 2;
            const int r0 = je + 4 * ti, c0 = je + 4 * tj;
            T acc[4][4];
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[x][y] = (T)0;
            for (int k = 0; k < pw; k += 4) {
                T pr[4][4], pc[4][4];
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const bool kk = k + q < pw;
                        pr[x][q] = (kk && r0 + x < kb) ? S[pk(r0 + x, jb + k + q)] : (T)0;
                        pc[x][q] = (kk && c0 + x < kb) ? S[pk(c0 + x, jb + k + q)] : (T)0;
                    }
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y)
#pragma unroll
                        for (int q = 0; q < 4; ++q) acc[x][y] += pr[x][q] * pc[y][q];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    const int r = r0 + x, c = c0 + y;
                    if (r < kb && c <= r) S[pk(r, c)] -= acc[x][y];
                }
        }
        __syncthreads();
    }
    for (int i = wave; i < kb; i += NT / 64)
        for (int j = lane; j <= i; j += 64) D[(long)i * lda + j] = S[pk(i, j)];
}

// L21 = A21 L11^-T for rows [r0, r0 + nrows): each row x solves x L11^T = a,
// i.e. x_j = (a_j - sum_{q<j} x_q L11[j][q]) / L11[j][j].
template <typename T>
__global__ __launch_bounds__(NT) void trsm_rltn_kernel(T *A, long lda, int k0, int kb, int r0, int nrows,
                                                       const int *info)
{
    if (*info != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *Ls = reinterpret_cast<T *>(smem_raw);        // packed lower incl. diagonal: row i at i*(i+1)/2
    T *Bs = Ls + (size_t)kb * (kb + 1) / 2;         // kb x 65 (transposed strip: Bs[j*65 + row])
    const int t = threadIdx.x, lane = t & 63, rg = t >> 6;
    const T *L = A + (long)k0 * lda + k0;
    for (int i = rg; i < kb; i += NT / 64) {
        T *dst = Ls + (size_t)i * (i + 1) / 2;
        for (int j = lane; j <= i; j += 64) dst[j] = L[(long)i * lda + j];
    }
    const int rbase = r0 + blockIdx.x * 64;
    {
        // the 64 x kb row strip, transposed into Bs[j*65 + r], 8 loads in flight per thread
        const int vrows = min(64, r0 + nrows - rbase);
        constexpr int U = 8;
        for (int b = threadIdx.x; b < 64 * kb; b += U * NT) {
            T v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = b + u * NT, r = idx / kb, j = idx - r * kb;
                v[u] = (idx < 64 * kb && r < vrows) ? A[(long)(rbase + r) * lda + k0 + j] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = b + u * NT, r = idx / kb, j = idx - r * kb;
                if (idx < 64 * kb) Bs[j * 65 + r] = v[u];
            }
        }
    }
    __syncthreads();
    for (int jb = 0; jb < kb; jb += 8) {
        const int je = min(jb + 8, kb);
        if (rg == 0) {
            // the 8 columns of the block in registers: x_j = (b_j - sum_q L[j][q] x_q) / L[j][j]
            T x[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) x[q] = (jb + q < je) ? Bs[(jb + q) * 65 + lane] : (T)0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (jb + q < je) {
                    const T *Lj = Ls + (size_t)(jb + q) * (jb + q + 1) / 2 + jb;
#pragma unroll
                    for (int pp = 0; pp < q; ++pp) x[q] -= Lj[pp] * x[pp];
                    x[q] = x[q] / Lj[q];
                    Bs[(jb + q) * 65 + lane] = x[q];
                }
            }
        }
        __syncthreads();
        T x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = (jb + q < je) ? Bs[(jb + q) * 65 + lane] : (T)0;
        // later columns, four per wave at a time (all loads issued before the arithmetic)
        for (int j0 = je + rg; j0 < kb; j0 += 4 * (NT / 64)) {
            T acc[4], lv[4][8];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = min(j0 + u * (NT / 64), kb - 1);
                acc[u] = Bs[j * 65 + lane];
                const T *Lj = Ls + (size_t)j * (j + 1) / 2 + jb;
#pragma unroll
                for (int q = 0; q < 8; ++q) lv[u][q] = (jb + q < je) ? Lj[q] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < 8; ++q) acc[u] -= lv[u][q] * x[q];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (j0 + u * (NT / 64) < kb) Bs[(j0 + u * (NT / 64)) * 65 + lane] = acc[u];
        }
        __syncthreads();
    }
    for (int r = rg; r < 64; r += NT / 64) {
        const int gr = rbase + r;
        if (gr < r0 + nrows)
            for (int j = lane; j < kb; j += 64) A[(long)gr * lda + k0 + j] = Bs[j * 65 + r];
    }
}

// Linv = L11^-1 of the 128 x 128 FP64 diagonal block (lower, non-unit), for
// the TRSM as a product L21 = A21 Linv^T.  Blocks of 32:
//   1. wave w inverts diagonal block w, lane c computing column c by forward
//      substitution in registers (true division by the diagonal);
//   2. block rows i = 1..3 in order, for each j < i:
//      T = sum_{k=j}^{i-1} L_ik Linv_kj, then Linv_ij = -Linv_ii T
//      (all 256 threads, one row x 4 columns of the 32 x 32 block each);
//   3. Linv (lower, zeros above) to global, row stride 128.
// The inverse's ten lower blocks and one T block live in LDS (88 KiB, so the
// kernel fits on a CU beside one workgroup of the trailing update running on
// the other stream); L is read from global memory (cached).
constexpr int TI64 = 128, TB64 = 32;
__device__ __forceinline__ int blk64(int i, int j) { return (i * (i + 1) / 2 + j) * TB64 * TB64; }
__global__ __launch_bounds__(NT) void trtri_lower64_kernel(const double *L, long ldl, double *Linv, const int *info)
{
    if (*info != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *X = reinterpret_cast<double *>(smem_raw);   // 10 blocks of 32 x 32
    double *Ts = X + 10 * TB64 * TB64;                   // 32 x 32
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    {
        const double *D = L + (long)(TB64 * wave) * ldl + TB64 * wave;
        const int c = lane;
        double x[TB64];
#pragma unroll
        for (int i = 0; i < TB64; ++i) {
            double sacc = (i == c) ? 1.0 : 0.0;
#pragma unroll
            for (int j = 0; j < i; ++j) sacc -= D[(long)i * ldl + j] * x[j];
            x[i] = sacc / D[(long)i * ldl + i];
        }
        if (c < TB64)
#pragma unroll
            for (int i = 0; i < TB64; ++i) X[blk64(wave, wave) + i * TB64 + c] = x[i];
    }
    __syncthreads();
    const int r = t >> 3, c4 = t & 7;   // this thread's row and four columns of a block
    for (int i = 1; i < TI64 / TB64; ++i) {
        const double *Lr = L + (long)(TB64 * i + r) * ldl;
        for (int j = 0; j < i; ++j) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            for (int k = j; k < i; ++k) {
                const double *xk = X + blk64(k, j) + 4 * c4;
#pragma unroll 8
                for (int q = 0; q < TB64; ++q) {
                    const double a = Lr[TB64 * k + q];
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[e] = fma(a, xk[q * TB64 + e], acc[e]);
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) Ts[r * TB64 + 4 * c4 + e] = acc[e];
            __syncthreads();
            const double *dr = X + blk64(i, i) + r * TB64;
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] = 0.0;
            for (int q = 0; q <= r; ++q) {
                const double a = dr[q];
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[e] = fma(a, Ts[q * TB64 + 4 * c4 + e], acc[e]);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) X[blk64(i, j) + r * TB64 + 4 * c4 + e] = -acc[e];
            __syncthreads();
        }
    }
    for (int e = t; e < TI64 * TI64; e += NT) {
        const int rr = e / TI64, cc = e % TI64;
        Linv[e] = cc <= rr ? X[blk64(rr / TB64, cc / TB64) + (rr % TB64) * TB64 + cc % TB64] : 0.0;
    }
}

template <typename T>
__global__ void zero_upper_kernel(T *A, long lda, int n)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * n) return;
    int i = (int)(idx / n), j = (int)(idx % n);
    if (j > i) A[(long)i * lda + j] = (T)0;
}
} // namespace

// Look-ahead as in the LU: after block k's TRSM the main stream updates only
// the next block column (rows below it) and goes on to block k+1; the rest of
// the trailing lower triangle is updated on the side stream meanwhile.  Block
// k's side-stream update also covers block k+2's columns, so block k+1's
// next-column update waits for it.
template <typename T>
int scirs_potrf_device(int n, T *A, long lda, int kb, ScirsWork &w, hipStream_t stream)
{
    if (n <= 0 || kb <= 0 || kb > 128) return SCIRS_ERR_ARG;
    const bool lookahead = w.lookahead && w.side_stream && n > 2 * kb;
    if (lookahead) {
        SCIRS_HIP_CHECK(hipEventRecord(w.ev_start, stream));
        SCIRS_HIP_CHECK(hipStreamWaitEvent(w.side_stream, w.ev_start, 0));
    }
    for (int k0 = 0, p = 0; k0 < n; k0 += kb, ++p) {
        const int kbw = std::min(kb, n - k0);
        const size_t lds_p = (size_t)kbw * (kbw + 1) / 2 * sizeof(T);
        if (int rc = scirs_allow_lds((const void *)potf2_kernel<T>, lds_p)) return rc;
        hipLaunchKernelGGL(potf2_kernel<T>, dim3(1), dim3(NT), lds_p, stream,
                           A, lda, k0, kbw, w.info, w.info_val);
        const int rest = n - k0 - kbw;
        if (rest > 0) {
            static const bool inv_trsm = [] {
                const char *v = getenv("SCIRS_GPU_CHOL_INV_TRSM");
                return !(v && *v && atoi(v) == 0);
            }();
            if (std::is_same<T, double>::value && kbw == TI64 && w.linv64 && inv_trsm) {
                // L21 = A21 L11^-T as a product on the matrix cores: L11^-1 by one
                // workgroup, then one 128-row tile of A21 per workgroup, in place
                const size_t lds_i = (size_t)11 * TB64 * TB64 * sizeof(double);
                if (int rc = scirs_allow_lds((const void *)trtri_lower64_kernel, lds_i)) return rc;
                hipLaunchKernelGGL(trtri_lower64_kernel, dim3(1), dim3(NT), lds_i, stream,
                                   (const double *)(A + (long)k0 * lda + k0), lda, w.linv64, (const int *)w.info);
                SCIRS_HIP_CHECK(hipGetLastError());
                double *A21 = (double *)(A + (long)(k0 + kbw) * lda + k0);
                if (int rc = scirs_launch_dgemm_assign_t(rest, TI64, TI64, A21, lda, w.linv64, TI64, A21, lda,
                                                         (const int *)w.info, stream))
                    return rc;
            } else {
                const size_t lds = ((size_t)kbw * (kbw + 1) / 2 + (size_t)kbw * 65) * sizeof(T);
                if (int rc = scirs_allow_lds((const void *)trsm_rltn_kernel<T>, lds)) return rc;
                hipLaunchKernelGGL(trsm_rltn_kernel<T>, dim3((rest + 63) / 64), dim3(NT), lds, stream,
                                   A, lda, k0, kbw, k0 + kbw, rest, (const int *)w.info);
            }
            const T *L21 = A + (long)(k0 + kbw) * lda + k0;
            T *A22 = A + (long)(k0 + kbw) * lda + k0 + kbw;
            const int nn = lookahead ? std::min(kb, rest) : rest;   // columns updated on the main stream
            if (lookahead) {
                // the previous block's side-stream update covers these columns
                if (p > 0) SCIRS_HIP_CHECK(hipStreamWaitEvent(stream, w.ev_upd, 0));
                if (rest > nn) {
                    SCIRS_HIP_CHECK(hipEventRecord(w.ev_trsm, stream));
                    SCIRS_HIP_CHECK(hipStreamWaitEvent(w.side_stream, w.ev_trsm, 0));
                    int rc = scirs_syrk_update<T>(rest - nn, kbw, L21 + (long)nn * lda, lda,
                                                  A22 + (long)nn * lda + nn, lda, w, w.side_stream);
                    if (rc) return rc;
                }
                SCIRS_HIP_CHECK(hipEventRecord(w.ev_upd, w.side_stream));
                // A22[:, 0:nn) for every row below the block, in one launch (the strict
                // upper part of the nn x nn corner is updated too: it is never read,
                // and the factorisation zeroes it at the end)
                int rc = scirs_launch_gemm_sub_t<T>(rest, nn, kbw, L21, lda, L21, lda, A22, lda, w, stream);
                if (rc) return rc;
            } else {
                int rc = scirs_syrk_update<T>(rest, kbw, L21, lda, A22, lda, w, stream);
                if (rc) return rc;
            }
        }
        SCIRS_HIP_CHECK(hipGetLastError());
    }
    if (lookahead) {
        SCIRS_HIP_CHECK(hipEventRecord(w.ev_upd, w.side_stream));
        SCIRS_HIP_CHECK(hipStreamWaitEvent(stream, w.ev_upd, 0));
    }
    long total = (long)n * n;
    hipLaunchKernelGGL(zero_upper_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, A, lda, n);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

template int scirs_potrf_device<double>(int, double *, long, int, ScirsWork &, hipStream_t);
template int scirs_potrf_device<float>(int, float *, long, int, ScirsWork &, hipStream_t);
