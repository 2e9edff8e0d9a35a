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
#include <algorithm>

namespace {

constexpr int NT = 256;  // threads per workgroup everywhere in this file

template <typename T>
struct PanelArgs {
    T *A; long lda;
    int k0, m, kb, R, G;
    int64_t *ipiv_unused;
    T *cval; int *cidx; T *crow; T *rowj;
    unsigned *bar; unsigned *done; int *abort;
    int *pairs; int *npairs;
    int *info; double *info_val;
    unsigned spin_limit;
};

template <typename T>
__global__ __launch_bounds__(NT) void getf2_panel_kernel(PanelArgs<T> a)
{
    if (*a.info != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int kb = a.kb;
    const int kbp = kb + 1;                       // LDS pitch (breaks bank aliasing)
    T *tile = reinterpret_cast<T *>(smem_raw);    // R x kbp
    T *prow = tile + (size_t)a.R * kbp;           // kb
    T *lcol = prow + kb;                          // R multipliers of the current column
    int *top_label = reinterpret_cast<int *>(lcol + a.R);  // kb   (workgroup 0 only)
    int *out_pos = top_label + kb;                // kb
    int *out_label = out_pos + kb;                // kb
    __shared__ T red_v[NT / 64];
    __shared__ int red_i[NT / 64];
    __shared__ int s_p, s_g, s_nout, s_stop;
    __shared__ T s_pv;

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int g = blockIdx.x;
    const int row0 = g * a.R;                     // first panel-relative row of this chunk
    const int Rg = max(0, min(a.R, a.m - row0));  // rows held here
    T *Ap = a.A + (long)a.k0 * a.lda + a.k0;      // panel origin

    for (int r = wave; r < Rg; r += NT / 64)
        for (int c = lane; c < kb; c += 64)
            tile[r * kbp + c] = Ap[(long)(row0 + r) * a.lda + c];
    if (g == 0) {
        for (int q = t; q < kb; q += NT) top_label[q] = q;
        if (t == 0) s_nout = 0;
    }
    if (t == 0) s_stop = 0;
    __syncthreads();

    const T eps = num_limits<T>::eps();
    int buf = 0;
    for (int j = 0; j < kb; ++j) {
        // ---- 1. local arg-max over rows with panel index >= j
        const int rstart = max(0, j - row0);
        T bv = (T)0; int bi = -1;
        for (int r = rstart + t; r < Rg; r += NT) {
            T v = fabs(tile[r * kbp + j]);
            int i = row0 + r;
            if (bi < 0 || v > bv) { bv = v; bi = i; }   // rows visited in increasing order
        }
        wave_argmax(bv, bi);
        if (lane == 0) { red_v[wave] = bv; red_i[wave] = bi; }
        __syncthreads();
        if (t == 0) {
            T v = red_v[0]; int i = red_i[0];
            for (int w = 1; w < NT / 64; ++w) {
                int oi = red_i[w]; T ov = red_v[w];
                if (oi >= 0 && (i < 0 || ov > v || (ov == v && oi < i))) { v = ov; i = oi; }
            }
            s_p = i; s_pv = v;
        }
        __syncthreads();
        // ---- 2. publish the candidate row, and row j when this chunk holds it
        T *cv = a.cval + buf * a.G;
        int *ci = a.cidx + buf * a.G;
        T *cr = a.crow + ((size_t)buf * a.G + g) * kb;
        T *rj = a.rowj + (size_t)buf * kb;
        const int my_best = s_p;
        if (t == 0) { cv[g] = s_pv; ci[g] = my_best; }
        if (my_best >= 0)
            for (int c = t; c < kb; c += NT) cr[c] = tile[(my_best - row0) * kbp + c];
        if (j >= row0 && j < row0 + Rg)
            for (int c = t; c < kb; c += NT) rj[c] = tile[(j - row0) * kbp + c];
        // ---- 3. one grid barrier per column
        if (!grid_barrier(a.bar, (unsigned)(j + 1) * a.G, a.abort, a.spin_limit)) break;
        // ---- 4. every workgroup reduces the same candidate list
        {
            T v = (T)0; int i = -1, wg = -1;
            for (int q = t; q < a.G; q += NT) {
                int qi = ci[q]; T qv = cv[q];
                if (qi >= 0 && (i < 0 || qv > v || (qv == v && qi < i))) { v = qv; i = qi; wg = q; }
            }
            // arg-max carrying the owning workgroup: pack (row, wg) — rows are unique
            T vv = v; int ii = i;
            wave_argmax(vv, ii);
            // recover wg of the winning row: the lane that owned it broadcasts
            unsigned long long ball = __ballot(i == ii && i >= 0);
            int src = ball ? __ffsll(ball) - 1 : 0;
            int wgw = __shfl(wg, src, 64);
            if (lane == 0) { red_v[wave] = vv; red_i[wave] = ii; }
            __shared__ int red_g[NT / 64];
            if (lane == 0) red_g[wave] = wgw;
            __syncthreads();
            if (t == 0) {
                T bv2 = red_v[0]; int bi2 = red_i[0], bg2 = red_g[0];
                for (int w = 1; w < NT / 64; ++w) {
                    int oi = red_i[w]; T ov = red_v[w];
                    if (oi >= 0 && (bi2 < 0 || ov > bv2 || (ov == bv2 && oi < bi2))) {
                        bv2 = ov; bi2 = oi; bg2 = red_g[w];
                    }
                }
                s_p = bi2; s_g = bg2; s_pv = bv2;
                if (!(bv2 >= eps)) {                 // |pivot| < eps: singular (lapack.rs:106-120)
                    s_stop = 1;
                    if (g == 0) {
                        *a.info = a.k0 + j + 1;
                        *a.info_val = (double)bv2;
                    }
                }
            }
            __syncthreads();
        }
        if (s_stop) break;
        const int p = s_p;
        const T *wrow = a.crow + ((size_t)buf * a.G + s_g) * kb;
        for (int c = t; c < kb; c += NT) prow[c] = wrow[c];
        __syncthreads();
        // ---- 5. row interchange inside the panel
        if (p != j) {
            if (j >= row0 && j < row0 + Rg)
                for (int c = t; c < kb; c += NT) tile[(j - row0) * kbp + c] = prow[c];
            if (p >= row0 && p < row0 + Rg)
                for (int c = t; c < kb; c += NT) tile[(p - row0) * kbp + c] = rj[c];
        }
        if (g == 0 && wave == 0) {
            // bookkeeping of the net permutation (panel-relative positions)
            if (p != j) {
                if (p < kb) {
                    if (lane == 0) {
                        int tmp = top_label[j]; top_label[j] = top_label[p]; top_label[p] = tmp;
                    }
                } else {
                    int nout = s_nout, found = -1;
                    for (int base = 0; base < nout && found < 0; base += 64) {
                        int s = base + lane;
                        unsigned long long b = __ballot(s < nout && out_pos[s] == p);
                        if (b) found = base + __ffsll(b) - 1;
                    }
                    if (lane == 0) {
                        if (found < 0) {
                            found = nout;
                            out_pos[found] = p;
                            out_label[found] = p;
                            s_nout = nout + 1;
                        }
                        int tmp = top_label[j]; top_label[j] = out_label[found]; out_label[found] = tmp;
                    }
                }
            }
        }
        __syncthreads();
        // ---- 6. multipliers (true division, lapack.rs:138) and rank-1 update
        const T pv = prow[j];
        const int r1 = max(0, j + 1 - row0);
        for (int r = r1 + t; r < Rg; r += NT) {
            T l = tile[r * kbp + j] / pv;
            tile[r * kbp + j] = l;
            lcol[r] = l;
        }
        __syncthreads();
        const int ncol = kb - j - 1;
        if (ncol > 0 && Rg > r1) {
            // 2-D split: consecutive threads walk consecutive columns of one row
            const int nrow = Rg - r1;
            const long total = (long)nrow * ncol;
            for (long w = t; w < total; w += NT) {
                int r = r1 + (int)(w / ncol);
                int c = j + 1 + (int)(w % ncol);
                tile[r * kbp + c] -= lcol[r] * prow[c];
            }
        }
        __syncthreads();
        buf ^= 1;
    }

    // write the factored panel back
    for (int r = wave; r < Rg; r += NT / 64)
        for (int c = lane; c < kb; c += 64)
            Ap[(long)(row0 + r) * a.lda + c] = tile[r * kbp + c];

    if (g == 0) {
        // gather list (dst <- src) of the net permutation, absolute rows
        __syncthreads();
        if (t == 0) {
            int np = 0;
            for (int q = 0; q < kb; ++q)
                if (top_label[q] != q) {
                    a.pairs[2 * np] = a.k0 + q; a.pairs[2 * np + 1] = a.k0 + top_label[q]; ++np;
                }
            for (int s = 0; s < s_nout; ++s)
                if (out_label[s] != out_pos[s]) {
                    a.pairs[2 * np] = a.k0 + out_pos[s]; a.pairs[2 * np + 1] = a.k0 + out_label[s]; ++np;
                }
            *a.npairs = np;
        }
    }

    // the last workgroup out resets the barrier state for the next launch
    __syncthreads();
    if (t == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        unsigned prev = __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == (unsigned)a.G - 1) {
            __hip_atomic_store(a.bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(a.done, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

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
    for (int q = wave; q < np; q += NT / 64)
        if (ok) vals[q * 64 + lane] = A[(long)pairs[2 * q + 1] * lda + c];
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
    for (int i = 1 + rg; i < kb; i += NT / 64) {
        T *dst = Ls + (size_t)i * (i - 1) / 2;
        for (int j = lane; j < i; j += 64) dst[j] = L[(long)i * lda + j];
    }
    for (int i = rg; i < kb; i += NT / 64)
        Bs[i * 64 + lane] = ok ? A[(long)(k0 + i) * lda + c] : (T)0;
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

size_t panel_lds_bytes(int R, int kb, size_t esz)
{
    return ((size_t)R * (kb + 1) + kb + R) * esz + 3 * (size_t)kb * sizeof(int);
}

template <typename T>
int max_lds_rows(int kb)
{
    const size_t budget = 160 * 1024 - 2048;   // leave room for the static reduction buffers
    size_t per_row = (size_t)(kb + 2) * sizeof(T);
    size_t fixed = (size_t)kb * sizeof(T) + 3 * (size_t)kb * sizeof(int);
    return (int)((budget - fixed) / per_row);
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
    const int rmax = max_lds_rows<T>(kb);
    for (int p = 0; p < npanels; ++p) {
        const int k0 = p * kb;
        const int kbw = std::min(kb, kmax - k0);
        const int mrows = m - k0;
        int G = (mrows + rmax - 1) / rmax;
        // spread rows over more workgroups when there is room: fewer rows per
        // workgroup shortens every column step
        const int target = std::max(8, w.panel_rows);
        G = std::max(G, std::min(w.max_g, (mrows + target - 1) / target));
        G = std::min(G, w.max_g);
        const int R = (mrows + G - 1) / G;
        if (R > rmax) {
            scirs_set_error("getrf: panel of %d rows does not fit %d co-resident workgroups", mrows, w.max_g);
            return SCIRS_ERR_UNSUPPORTED;
        }
        G = (mrows + R - 1) / R;
        PanelArgs<T> pa;
        pa.A = A; pa.lda = lda; pa.k0 = k0; pa.m = mrows; pa.kb = kbw; pa.R = R; pa.G = G;
        pa.ipiv_unused = nullptr;
        pa.cval = (T *)w.cval; pa.cidx = w.cidx; pa.crow = (T *)w.crow; pa.rowj = (T *)w.rowj;
        pa.bar = w.bar; pa.done = w.done; pa.abort = w.abort;
        pa.pairs = w.pairs + (size_t)p * 4 * SCIRS_KB_MAX; pa.npairs = w.npairs + p;
        pa.info = w.info; pa.info_val = w.info_val; pa.spin_limit = w.spin_limit;
        void *args[] = {&pa};
        const size_t lds = panel_lds_bytes(R, kbw, sizeof(T));
        if (int rc = scirs_allow_lds((const void *)getf2_panel_kernel<T>, lds)) return rc;
        SCIRS_HIP_CHECK(hipLaunchCooperativeKernel((const void *)getf2_panel_kernel<T>, dim3(G), dim3(NT),
                                                   args, lds, stream));
        // row interchanges outside the panel
        const int nout_cols = n - kbw;
        if (nout_cols > 0) {
            const size_t lds_sw = (size_t)2 * kbw * 64 * sizeof(T);
            if (int rc = scirs_allow_lds((const void *)laswp_kernel<T>, lds_sw)) return rc;
            hipLaunchKernelGGL(laswp_kernel<T>, dim3((nout_cols + 63) / 64), dim3(NT), lds_sw, stream,
                               A, lda, n, k0, kbw, pa.pairs, pa.npairs, (const int *)w.info);
            SCIRS_HIP_CHECK(hipGetLastError());
        }
        const int nright = n - k0 - kbw;
        if (nright > 0) {
            const size_t lds_t = ((size_t)kbw * (kbw - 1) / 2 + (size_t)kbw * 64) * sizeof(T);
            if (int rc = scirs_allow_lds((const void *)trsm_llnu_kernel<T>, lds_t)) return rc;
            hipLaunchKernelGGL(trsm_llnu_kernel<T>, dim3((nright + 63) / 64), dim3(NT), lds_t, stream,
                               A, lda, k0, kbw, n, (const int *)w.info);
            SCIRS_HIP_CHECK(hipGetLastError());
            const int mbelow = m - k0 - kbw;
            if (mbelow > 0) {
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
