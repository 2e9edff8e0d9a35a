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

namespace {
constexpr int NT = 256;

template <typename T>
__global__ __launch_bounds__(NT) void potf2_kernel(T *A, long lda, int k0, int kb, int *info, double *info_val)
{
    if (*info != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *S = reinterpret_cast<T *>(smem_raw);   // kb x (kb+1)
    const int p = kb + 1;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    T *D = A + (long)k0 * lda + k0;
    g2s_block<T>(S, p, D, lda, kb, kb, kb, kb);   // (the strict upper part is never read)
    __shared__ int s_fail;
    if (t == 0) s_fail = 0;
    __syncthreads();
    for (int i = 0; i < kb; ++i) {
        if (t == 0) {
            T v = S[i * p + i];
            if (!(v > (T)0)) {
                s_fail = 1;
                *info = k0 + i + 1;
                *info_val = (double)v;
            } else {
                S[i * p + i] = sqrt(v);
            }
        }
        __syncthreads();
        if (s_fail) return;
        const T d = S[i * p + i];
        for (int r = i + 1 + t; r < kb; r += NT) S[r * p + i] = S[r * p + i] / d;
        __syncthreads();
        // trailing lower update of the block: S[r][c] -= S[r][i] * S[c][i], i < c <= r
        for (int r = i + 1 + wave; r < kb; r += NT / 64) {
            const T lri = S[r * p + i];
            for (int c = i + 1 + lane; c <= r; c += 64) S[r * p + c] -= lri * S[c * p + i];
        }
        __syncthreads();
    }
    for (int i = wave; i < kb; i += NT / 64)
        for (int j = lane; j <= i; j += 64) D[(long)i * lda + j] = S[i * p + j];
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
            for (int j = jb; j < je; ++j) {
                T s = Bs[j * 65 + lane];
                const T *Lj = Ls + (size_t)j * (j + 1) / 2;
                for (int q = jb; q < j; ++q) s -= Lj[q] * Bs[q * 65 + lane];
                Bs[j * 65 + lane] = s / Lj[j];
            }
        }
        __syncthreads();
        T x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = (jb + q < je) ? Bs[(jb + q) * 65 + lane] : (T)0;
        for (int j = je + rg; j < kb; j += NT / 64) {
            const T *Lj = Ls + (size_t)j * (j + 1) / 2 + jb;
            T s = Bs[j * 65 + lane];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (jb + q < je) s -= Lj[q] * x[q];
            Bs[j * 65 + lane] = s;
        }
        __syncthreads();
    }
    for (int r = rg; r < 64; r += NT / 64) {
        const int gr = rbase + r;
        if (gr < r0 + nrows)
            for (int j = lane; j < kb; j += 64) A[(long)gr * lda + k0 + j] = Bs[j * 65 + r];
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

template <typename T>
int scirs_potrf_device(int n, T *A, long lda, int kb, ScirsWork &w, hipStream_t stream)
{
    if (n <= 0 || kb <= 0 || kb > 128) return SCIRS_ERR_ARG;
    for (int k0 = 0; k0 < n; k0 += kb) {
        const int kbw = std::min(kb, n - k0);
        if (int rc = scirs_allow_lds((const void *)potf2_kernel<T>, (size_t)kbw * (kbw + 1) * sizeof(T))) return rc;
        hipLaunchKernelGGL(potf2_kernel<T>, dim3(1), dim3(NT), (size_t)kbw * (kbw + 1) * sizeof(T), stream,
                           A, lda, k0, kbw, w.info, w.info_val);
        const int rest = n - k0 - kbw;
        if (rest > 0) {
            const size_t lds = ((size_t)kbw * (kbw + 1) / 2 + (size_t)kbw * 65) * sizeof(T);
            if (int rc = scirs_allow_lds((const void *)trsm_rltn_kernel<T>, lds)) return rc;
            hipLaunchKernelGGL(trsm_rltn_kernel<T>, dim3((rest + 63) / 64), dim3(NT), lds, stream,
                               A, lda, k0, kbw, k0 + kbw, rest, (const int *)w.info);
            int rc = scirs_syrk_update<T>(rest, kbw, A + (long)(k0 + kbw) * lda + k0, lda,
                                          A + (long)(k0 + kbw) * lda + k0 + kbw, lda, w, stream);
            if (rc) return rc;
        }
        SCIRS_HIP_CHECK(hipGetLastError());
    }
    long total = (long)n * n;
    hipLaunchKernelGGL(zero_upper_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, A, lda, n);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

template int scirs_potrf_device<double>(int, double *, long, int, ScirsWork &, hipStream_t);
template int scirs_potrf_device<float>(int, float *, long, int, ScirsWork &, hipStream_t);
