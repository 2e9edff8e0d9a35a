// Blocked right-looking LU with partial pivoting for gfx950.
//
// Algorithm (row-major A, the layout of an ndarray in standard order), per
// panel of kb columns starting at k0:
//   1. getf2_panel  — cooperative kernel.  The panel's m-k0 rows are split
//      into contiguous chunks, one per workgroup, and kept in LDS for the whole
//      panel.  Each column: wave64 arg-max over the chunk, candidates published
//      to global memory, one grid barrier, every workgroup reduces the same
//      candidate list (lowest row wins ties, the strict '>' scan of
//      scirs2-linalg/src/lapack.rs:95-104), swaps, divides and applies the
//      rank-1 update to its own rows.  The absolute singularity test of
//      lapack.rs:106-120 (|pivot| < eps) stops the factorisation with
//      info = column + 1.
//   2. laswp        — applies the panel's net row permutation to the columns
//      left and right of the panel as one gather (all loads, then all stores).
//   3. trsm         — U12 = L11^-1 A12, 64-column strips with L11 (packed) and
//      the strip in LDS.
//   4. dgemm        — A22 -= L21 U12 on the matrix cores (gemm_f64.hip).
//
// The arithmetic per element is the reference's (multiplier by true division,
// a_ij -= l_ik * u_kj); only the order of the k-sum differs (blocked), which
// is what the normwise backward-error parity test allows.
#include "common.hpp"
#include "kernels.hpp"
#include "panel.hpp"
#include "prof.hpp"
#include <algorithm>

namespace {

constexpr int NT = 256;  // threads per workgroup everywhere in this file


// Apply a gather list to the columns outside [k0, k0+kb): each workgroup owns 64
// columns, loads every source row of its columns first, then stores.
template <typename T>
__global__ __launch_bounds__(NT) void laswp_kernel(T *A, long lda, int ncols_total, int k0, int kb,
                                                   const int *pairs, const int *npairs, const int *info)
{
    if (*info != 0) return;
    const int np = *npairs;
    if (np == 0) return;
    // column c of the workgroup: skip the panel columns
    int c = blockIdx.x * 64 + (threadIdx.x & 63);
    if (c >= k0) c += kb;
    const bool ok = c < ncols_total;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *vals = reinterpret_cast<T *>(smem_raw);   // np x 64
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int U = 8;   // loads in flight per thread before the LDS stores
    for (int q0 = wave; q0 < np; q0 += U * (NT / 64)) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = q0 + u * (NT / 64);
            v[u] = (ok && q < np) ? A[(long)pairs[2 * q + 1] * lda + c] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = q0 + u * (NT / 64);
            if (q < np) vals[q * 64 + lane] = v[u];
        }
    }
    __syncthreads();
    for (int q = wave; q < np; q += NT / 64)
        if (ok) A[(long)pairs[2 * q] * lda + c] = vals[q * 64 + lane];
}

// U12 = L11^-1 A12 for the kb x (N - k0 - kb) block right of the panel.
// One workgroup per 64-column strip; L11 (strict lower, packed by rows) and
// the strip live in LDS; rows are solved in blocks of 8.
template <typename T>
__global__ __launch_bounds__(NT) void trsm_llnu_kernel(T *A, long lda, int k0, int kb, int ncols_total,
                                                       const int *info)
{
    if (*info != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *Ls = reinterpret_cast<T *>(smem_raw);        // packed strict lower: row i at i*(i-1)/2
    T *Bs = Ls + (size_t)kb * (kb - 1) / 2;         // kb x 64
    const int t = threadIdx.x, lane = t & 63, rg = t >> 6;
    const int cbase = k0 + kb + blockIdx.x * 64;
    const int c = cbase + lane;
    const bool ok = c < ncols_total;
    const T *L = A + (long)k0 * lda + k0;
    // strict lower triangle of L11, packed by rows (row i at i*(i-1)/2), with
    // 8 loads in flight per thread
    {
        const int total = kb * (kb - 1) / 2;
        constexpr int U = 8;
        for (int b = threadIdx.x; b < total; b += U * NT) {
            T v[U]; int ii[U], jj[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = b + u * NT;
                // invert idx = i*(i-1)/2 + j, 0 <= j < i
                int i = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)idx)) * 0.5f);
                while (i * (i - 1) / 2 > idx) --i;
                while ((i + 1) * i / 2 <= idx) ++i;
                ii[u] = i; jj[u] = idx - i * (i - 1) / 2;
                v[u] = idx < total ? L[(long)i * lda + jj[u]] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (b + u * NT < total) Ls[(size_t)ii[u] * (ii[u] - 1) / 2 + jj[u]] = v[u];
        }
    }
    g2s_block<T>(Bs, 64, A + (long)k0 * lda + cbase, lda, kb, 64, kb, ncols_total - cbase);
    __syncthreads();
    for (int ib = 0; ib < kb; ib += 8) {
        const int ie = min(ib + 8, kb);
        if (rg == 0) {
            for (int i = ib; i < ie; ++i) {
                T s = Bs[i * 64 + lane];
                const T *Li = Ls + (size_t)i * (i - 1) / 2;
                for (int j = ib; j < i; ++j) s -= Li[j] * Bs[j * 64 + lane];
                Bs[i * 64 + lane] = s;
            }
        }
        __syncthreads();
        T x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = (ib + q < ie) ? Bs[(ib + q) * 64 + lane] : (T)0;
        for (int r = ie + rg; r < kb; r += NT / 64) {
            const T *Lr = Ls + (size_t)r * (r - 1) / 2 + ib;
            T s = Bs[r * 64 + lane];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (ib + q < ie) s -= Lr[q] * x[q];
            Bs[r * 64 + lane] = s;
        }
        __syncthreads();
    }
    if (ok)
        for (int i = rg; i < kb; i += NT / 64) A[(long)(k0 + i) * lda + c] = Bs[i * 64 + lane];
}

} // namespace

template <typename T>
int scirs_getrf_device(int m, int n, T *A, long lda, int kb, ScirsWork &w, hipStream_t stream)
{
    if (m <= 0 || n <= 0) return SCIRS_ERR_ARG;
    if (kb <= 0 || kb > SCIRS_KB_MAX) return SCIRS_ERR_ARG;
    const int kmax = std::min(m, n);
    const int npanels = (kmax + kb - 1) / kb;
    if (npanels > w.max_panels) {
        scirs_set_error("getrf: %d panels exceed the workspace (%d)", npanels, w.max_panels);
        return SCIRS_ERR_ARG;
    }
    const int rmax = panel::max_rows<T>(kb);
    for (int p = 0; p < npanels; ++p) {
        const int k0 = p * kb;
        const int kbw = std::min(kb, kmax - k0);
        const int mrows = m - k0;
        // rows per workgroup: as many as the target allows, never more than LDS holds
        const int target = std::max(8, std::min(rmax, w.panel_rows));
        int G = std::min(w.max_g, (mrows + target - 1) / target);
        G = std::max(G, 1);
        int R = (mrows + G - 1) / G;
        if (R > rmax) {
            G = (mrows + rmax - 1) / rmax;
            if (G > w.max_g) {
                scirs_set_error("getrf: panel of %d rows does not fit %d co-resident workgroups", mrows, w.max_g);
                return SCIRS_ERR_UNSUPPORTED;
            }
            R = (mrows + G - 1) / G;
        }
        G = (mrows + R - 1) / R;
        panel::PanelArgs<T> pa;
        pa.A = A; pa.lda = lda; pa.k0 = k0; pa.m = mrows; pa.kb = kbw; pa.R = R; pa.G = G;
        pa.xw = w.xw; pa.xrow = w.xrow; pa.xrj = w.xrj;
        pa.seq_base = w.seq_next;
        w.seq_next += 512u;
        pa.pairs = w.pairs + (size_t)p * 4 * SCIRS_KB_MAX; pa.npairs = w.npairs + p;
        pa.info = w.info; pa.info_val = w.info_val; pa.abort = w.abort; pa.spin_limit = w.spin_limit;
        void *args[] = {&pa};
        const size_t lds = panel::lds_bytes(R, kbw, sizeof(T));
        if (int rc = scirs_allow_lds((const void *)panel::getf2_panel_kernel<T>, lds)) return rc;
        {
            scirs::ProfScope ps(SCIRS_PROF_PANEL, (double)mrows * kbw * kbw, 2.0 * mrows * kbw * sizeof(T), stream);
            SCIRS_HIP_CHECK(hipLaunchCooperativeKernel((const void *)panel::getf2_panel_kernel<T>, dim3(G),
                                                       dim3(panel::NT), args, lds, stream));
        }
        // row interchanges outside the panel
        const int nout_cols = n - kbw;
        if (nout_cols > 0) {
            const size_t lds_sw = (size_t)2 * kbw * 64 * sizeof(T);
            if (int rc = scirs_allow_lds((const void *)laswp_kernel<T>, lds_sw)) return rc;
            scirs::ProfScope ps(SCIRS_PROF_LASWP, 0.0, 4.0 * kbw * nout_cols * sizeof(T), stream);
            hipLaunchKernelGGL(laswp_kernel<T>, dim3((nout_cols + 63) / 64), dim3(NT), lds_sw, stream,
                               A, lda, n, k0, kbw, pa.pairs, pa.npairs, (const int *)w.info);
            SCIRS_HIP_CHECK(hipGetLastError());
        }
        const int nright = n - k0 - kbw;
        if (nright > 0) {
            const size_t lds_t = ((size_t)kbw * (kbw - 1) / 2 + (size_t)kbw * 64) * sizeof(T);
            if (int rc = scirs_allow_lds((const void *)trsm_llnu_kernel<T>, lds_t)) return rc;
            {
                scirs::ProfScope ps(SCIRS_PROF_TRSM, (double)kbw * kbw * nright,
                                    (0.5 * kbw * kbw + 2.0 * kbw * nright) * sizeof(T), stream);
                hipLaunchKernelGGL(trsm_llnu_kernel<T>, dim3((nright + 63) / 64), dim3(NT), lds_t, stream,
                                   A, lda, k0, kbw, n, (const int *)w.info);
                SCIRS_HIP_CHECK(hipGetLastError());
            }
            const int mbelow = m - k0 - kbw;
            if (mbelow > 0) {
                scirs::ProfScope ps(SCIRS_PROF_UPDATE, 2.0 * mbelow * nright * kbw,
                                    (2.0 * mbelow * nright + (double)mbelow * kbw + (double)kbw * nright) * sizeof(T),
                                    stream);
                int rc = scirs_trailing_update<T>(mbelow, nright, kbw,
                                                  A + (long)(k0 + kbw) * lda + k0, lda,
                                                  A + (long)k0 * lda + k0 + kbw, lda,
                                                  A + (long)(k0 + kbw) * lda + k0 + kbw, lda,
                                                  w, stream);
                if (rc) return rc;
            }
        }
    }
    return SCIRS_OK;
}

// ---------------------------------------------------------------- permutation
namespace {
__global__ __launch_bounds__(NT) void perm_from_pairs_kernel(int m, int npanels, const int *pairs,
                                                             const int *npairs, int64_t *perm)
{
    for (int i = threadIdx.x; i < m; i += NT) perm[i] = i;
    __syncthreads();
    for (int p = 0; p < npanels; ++p) {
        const int *pp = pairs + (size_t)p * 4 * SCIRS_KB_MAX;
        const int np = npairs[p];
        int64_t v0 = 0, v1 = 0;
        const int q0 = threadIdx.x, q1 = threadIdx.x + NT;
        if (q0 < np) v0 = perm[pp[2 * q0 + 1]];
        if (q1 < np) v1 = perm[pp[2 * q1 + 1]];
        __syncthreads();
        if (q0 < np) perm[pp[2 * q0]] = v0;
        if (q1 < np) perm[pp[2 * q1]] = v1;
        __syncthreads();
    }
}

template <typename T>
__global__ void permute_rows_kernel(int n, int nrhs, const int64_t *perm, const T *B, long ldb, T *X, long ldx)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * nrhs) return;
    int i = (int)(idx / nrhs), j = (int)(idx % nrhs);
    X[(long)i * ldx + j] = B[perm[i] * ldb + j];
}
} // namespace

int scirs_perm_from_pairs(int m, int n, int kb, const ScirsWork &w, int64_t *perm, hipStream_t stream)
{
    const int npanels = (std::min(m, n) + kb - 1) / kb;
    hipLaunchKernelGGL(perm_from_pairs_kernel, dim3(1), dim3(NT), 0, stream, m, npanels, w.pairs, w.npairs, perm);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

template <typename T>
int scirs_permute_rows(int n, int nrhs, const int64_t *perm, const T *B, long ldb, T *X, long ldx,
                       hipStream_t stream)
{
    long total = (long)n * nrhs;
    hipLaunchKernelGGL(permute_rows_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                       n, nrhs, perm, B, ldb, X, ldx);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

template int scirs_getrf_device<double>(int, int, double *, long, int, ScirsWork &, hipStream_t);
template int scirs_getrf_device<float>(int, int, float *, long, int, ScirsWork &, hipStream_t);
template int scirs_permute_rows<double>(int, int, const int64_t *, const double *, long, double *, long, hipStream_t);
template int scirs_permute_rows<float>(int, int, const int64_t *, const float *, long, float *, long, hipStream_t);
