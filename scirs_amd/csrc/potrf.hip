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

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef double dv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));
template <typename T> struct MV;
template <> struct MV<double> {
    typedef dv4 V;
    static __device__ __forceinline__ V mfma(double a, double b, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of accumulator register q in lane group lk (column = lane % 16)
    static __device__ __forceinline__ int row(int lk, int q) { return lk + 4 * q; }
};
template <> struct MV<float> {
    typedef fv4 V;
    static __device__ __forceinline__ V mfma(float a, float b, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int q) { return 4 * lk + q; }
};

// The kb x kb diagonal block (kb <= 128) in LDS (lower triangle packed by
// rows), factored in column blocks of PB = 32:
//   A. wave 0 factors the PB x PB diagonal block, lane i holding row i in
//      registers: per column the diagonal by readlane, sqrt, true division of
//      the rows below (lapack.rs:661-676), the column handed to every lane
//      through LDS, and the later columns updated — one wave, no barrier;
//   B. the rows below solve x L_jj^T = a (one thread per row, left-looking,
//      true division, the diagonal block read from LDS as broadcasts);
//   C. the trailing lower triangle takes the rank-PB update on the matrix
//      cores (16 x 16 tiles, one wave per tile).
// The test of the reference, a_jj - sum <= 0 (lapack.rs:661-670), sets
// info = row + 1 and stops.
template <typename T>
__global__ __launch_bounds__(NT) void potf2_kernel(T *A, long lda, int k0, int kb, int *info, double *info_val)
{
    if (*info != 0) return;
    constexpr int PB = 32;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // the lower triangle, packed by rows (row r at r (r + 1) / 2): half the LDS
    // of a square tile, so the kernel fits on a CU beside one workgroup of the
    // trailing update running on the other stream
    T *S = reinterpret_cast<T *>(smem_raw);
    auto pk = [](int r, int c) { return r * (r + 1) / 2 + c; };
    __shared__ T colb[2][PB];
    // the factored diagonal block, unpacked, so that phase B addresses it with
    // one base register and constant offsets
    __shared__ T Dsq[PB][PB + 1];
    __shared__ int s_fail;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    T *D = A + (long)k0 * lda + k0;
    // rows wave, wave + 4, ...: 64 loads in flight per lane before any
    // LDS store (a load-then-store loop would wait for each row in turn)
    for (int i0 = wave; i0 < kb; i0 += 32 * (NT / 64)) {
        T v[32][2];
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int i = min(i0 + q * (NT / 64), kb - 1);
#pragma unroll
            for (int h = 0; h < 2; ++h) v[q][h] = D[(long)i * lda + min(lane + 64 * h, i)];
        }
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int i = i0 + q * (NT / 64);
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (i < kb && lane + 64 * h <= i) S[pk(i, lane + 64 * h)] = v[q][h];
        }
    }
    if (t == 0) s_fail = 0;
    __syncthreads();
    for (int jb = 0; jb < kb; jb += PB) {
        const int pw = min(PB, kb - jb), je = jb + pw;
        // A. the diagonal block
        if (wave == 0) {
            const int i = lane;
            T a[PB];
#pragma unroll
            for (int c = 0; c < PB; ++c) a[c] = (i < pw && c <= i) ? S[pk(jb + i, jb + c)] : (T)0;
            bool ok = true;
            int fail_row = -1;
            T fail_val = (T)0;
#pragma unroll
            for (int j = 0; j < PB; ++j) {
                if (ok && j < pw) {
                    const T dv = readlane_t<T>(a[j], j);
                    if (!(dv > (T)0)) {
                        ok = false;
                        fail_row = jb + j;
                        fail_val = dv;
                    } else {
                        const T d = sqrt(dv);
                        a[j] = (i == j) ? d : (i > j ? a[j] / d : a[j]);
                        T *cb = colb[j & 1];
                        if (i < PB) cb[i] = a[j];
                        wave_sync();
                        // rows at or above j take a zero multiplier instead of a
                        // select per column (a - 0 * l is a: l is finite)
                        const T lj = i > j ? a[j] : (T)0;
#pragma unroll
                        for (int c = j + 1; c < PB; ++c) a[c] -= lj * cb[c];
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < PB; ++c)
                if (i < pw && c <= i) S[pk(jb + i, jb + c)] = a[c];
            if (i < PB)
#pragma unroll
                for (int c = 0; c < PB; ++c) Dsq[i][c] = a[c];
            if (!ok && lane == 0) {
                s_fail = 1;
                *info = k0 + fail_row + 1;
                *info_val = (double)fail_val;
            }
        }
        __syncthreads();
        if (s_fail) return;
        if (je >= kb) break;
        // B. rows below: x L_jj^T = a, x_c = (a_c - sum_{q<c} x_q L[c][q]) / L[c][c]
        //    (left-looking, four partial sums to shorten the chain).  Every
        //    address is one base plus a constant: past pw the loads read the
        //    padding after the triangle or stale entries, and the values they
        //    give are masked before anything is stored.
        // (kb <= 128, so at most 96 rows: one per thread — a loop over rows
        // would let the compiler hoist the diagonal block's 528 loads out of it)
        if (je + t < kb) {
            const int r = je + t;
            T *xr = S + pk(r, jb);
            T x[PB];
#pragma unroll
            for (int c = 0; c < PB; ++c) {
                const T v = xr[c];
                x[c] = c < pw ? v : (T)0;
            }
#pragma unroll
            for (int c = 0; c < PB; ++c) {
                // row c of the block, addressed through an offset the compiler
                // sees only once x[c - 2] exists: its loads then overlap the
                // previous column's division, and 528 loads are never in flight
                // (or in registers) at once
                int off = c * (PB + 1);
                if (c >= 2) asm volatile("" : "+v"(off) : "v"(x[c - 2]));
                const T *drow = &Dsq[0][0] + off;
                T sp[4] = {x[c], (T)0, (T)0, (T)0};
#pragma unroll
                for (int q = 0; q < c; ++q) sp[q & 3] -= x[q] * drow[q];
                const T v = ((sp[0] + sp[1]) + (sp[2] + sp[3])) / drow[c];
                x[c] = c < pw ? v : (T)0;
            }
#pragma unroll
            for (int c = 0; c < PB; ++c)
                if (c < pw) xr[c] = x[c];
        }
        __syncthreads();
        // C. S[r][c] -= sum_k S[r][jb+k] S[c][jb+k], je <= c <= r < kb, 16 x 16 tiles
        {
            const int lr = lane & 15, lk = lane >> 4;
            const int nt = (kb - je + 15) / 16, ntiles = nt * (nt + 1) / 2;
            for (int tile = wave; tile < ntiles; tile += NT / 64) {
                int ti = 0;
                while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
                const int tj = tile - ti * (ti + 1) / 2;
                const int r0 = je + 16 * ti, c0 = je + 16 * tj;
                const int ra = min(r0 + lr, kb - 1), ca = min(c0 + lr, kb - 1);
                typename MV<T>::V acc = {0, 0, 0, 0};
                const T *ar = S + pk(ra, jb) + lk, *br = S + pk(ca, jb) + lk;
                T av[PB / 4], bv[PB / 4];
#pragma unroll
                for (int k = 0; k < PB / 4; ++k) {
                    const bool kk = 4 * k + lk < pw;
                    const T x = ar[4 * k], y = br[4 * k];
                    av[k] = kk ? x : (T)0;
                    bv[k] = kk ? y : (T)0;
                }
#pragma unroll
                for (int k = 0; k < PB / 4; ++k) acc = MV<T>::mfma(av[k], bv[k], acc);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = r0 + MV<T>::row(lk, q), c = c0 + lr;
                    if (r < kb && c <= r) S[pk(r, c)] -= acc[q];
                }
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
                // the row's 32 entries of L_ik in one round trip to memory
                double lv[TB64];
#pragma unroll
                for (int q = 0; q < TB64; ++q) lv[q] = Lr[TB64 * k + q];
#pragma unroll
                for (int q = 0; q < TB64; ++q)
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[e] = fma(lv[q], xk[q * TB64 + e], acc[e]);
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
        // the packed triangle and PB elements of padding that reads past a
        // partial column block may touch
        const size_t lds_p = ((size_t)kbw * (kbw + 1) / 2 + 32) * sizeof(T);
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
