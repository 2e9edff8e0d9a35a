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
//   3. trsm         — U12 = L11^-1 A12: FP64 by substitution in 64-column strips
//      with L11 (packed) and the strip in LDS; FP32 by forming L11^-1 in one
//      workgroup and multiplying the strips by it.
//   4. dgemm        — A22 -= L21 U12 on the matrix cores (gemm_f64.hip).
//
// The arithmetic per element is the reference's (multiplier by true division,
// a_ij -= l_ik * u_kj); only the order of the k-sum differs (blocked), which
// is what the normwise backward-error parity test allows.
#include "common.hpp"
#include "kernels.hpp"
#include "panel.hpp"
#include "prof.hpp"
#include "api_internal.hpp"
#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <type_traits>

namespace {

constexpr int NT = 256;  // threads per workgroup everywhere in this file


// Apply a gather list to nv columns: the first nleft are columns 0..nleft-1,
// the others start at column right0.  Each workgroup owns 64 columns, stages
// the (dst, src) list in LDS, loads every source row of its columns first (16
// loads in flight per thread), then stores.
template <typename T>
__global__ __launch_bounds__(NT) void laswp_kernel(T *A, long lda, int nleft, int right0, int nv,
                                                   const int *pairs, const int *npairs, const int *info)
{
    if (*info != 0) return;
    const int np = *npairs;
    if (np == 0) return;
    const int v = blockIdx.x * 64 + (threadIdx.x & 63);
    const int c = v < nleft ? v : right0 + (v - nleft);
    const bool ok = v < nv;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int *sp = reinterpret_cast<int *>(smem_raw);                   // 2 * np
    T *vals = reinterpret_cast<T *>(smem_raw + (size_t)4 * SCIRS_KB_MAX * sizeof(int));   // np x 64
    for (int q = threadIdx.x; q < 2 * np; q += NT) sp[q] = pairs[q];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int U = 16;   // loads in flight per thread before the LDS stores
    for (int q0 = wave; q0 < np; q0 += U * (NT / 64)) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = q0 + u * (NT / 64);
            v[u] = (ok && q < np) ? A[(long)sp[2 * q + 1] * lda + c] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = q0 + u * (NT / 64);
            if (q < np) vals[q * 64 + lane] = v[u];
        }
    }
    __syncthreads();
    for (int q = wave; q < np; q += NT / 64)
        if (ok) A[(long)sp[2 * q] * lda + c] = vals[q * 64 + lane];
}

// ---- FP32 TRSM through the inverse of L11 (the FP32 factorisation only
// preconditions refinement, so forming L11^-1 is acceptable there).
//
// trtri: one workgroup inverts the unit-lower kb x kb block (kb <= 128, padded
// with the identity to 128), in place in LDS: the 16 x 16 diagonal blocks by
// forward substitution (one lane per column, in registers), then three doubling
// levels
//     [A 0; B C]^-1 = [A^-1 0; -C^-1 B A^-1  C^-1]
// whose products run as 4 x 4 register tiles.  Result: L11^-1 stored
// TRANSPOSED (LinvT[k][r] = L11^-1[r][k], 128 x 128) in HBM.
constexpr int TI = 128;          // padded order of the inverted block
constexpr int TIP = TI + 1;      // LDS pitch of the block
constexpr int TP4 = TI + 4;      // LDS pitch of the transposed inverse (16-byte rows)

__global__ __launch_bounds__(NT) void trtri_unit_lower_kernel(const float *A, long lda, int kb, float *LinvT,
                                                              const int *info)
{
    if (*info != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *Ls = reinterpret_cast<float *>(smem_raw);   // TI x TIP: L, overwritten by its inverse
    float *Tm = Ls + TI * TIP;                         // 64 * 64: B A^-1 of one level
    const int t = threadIdx.x;
    g2s_block<float, 16>(Ls, TIP, A, lda, TI, TI, kb, kb);
    __syncthreads();
    for (int idx = t; idx < TI * TI; idx += NT) {
        const int i = idx / TI, j = idx % TI;
        if (j >= i) Ls[i * TIP + j] = (j == i) ? 1.0f : 0.0f;
    }
    __syncthreads();
    // 16 x 16 diagonal blocks: lane t < 128 -> block t / 16, column t % 16
    float x[16];
    const int b0 = (t / 16) * 16, c16 = t % 16;
    if (t < TI) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            float sum = 0.0f;
#pragma unroll
            for (int k = 0; k < i; ++k) sum += Ls[(b0 + i) * TIP + b0 + k] * x[k];
            x[i] = (i < c16) ? 0.0f : ((i == c16) ? 1.0f : -sum);
        }
    }
    __syncthreads();
    if (t < TI) {
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i > c16) Ls[(b0 + i) * TIP + b0 + c16] = x[i];
    }
    __syncthreads();
    for (int sz = 16; sz < TI; sz *= 2) {
        const int tpr = sz / 4, tiles = (TI / (2 * sz)) * tpr * tpr;
        // Tm[p] = B_p A_p^-1 with B_p = L[C rows, A cols], A_p^-1 = Ls[A, A]
        for (int tile = t; tile < tiles; tile += NT) {
            const int pidx = tile / (tpr * tpr), rem = tile % (tpr * tpr), ti = rem / tpr, tj = rem % tpr;
            const int a0 = 2 * pidx * sz, c0 = a0 + sz;
            float acc[4][4] = {};
            for (int k = 0; k < sz; ++k) {
                float bv[4], av[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    bv[u] = Ls[(c0 + 4 * ti + u) * TIP + a0 + k];
                    av[u] = Ls[(a0 + k) * TIP + a0 + 4 * tj + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] += bv[u] * av[v];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) Tm[pidx * sz * sz + (4 * ti + u) * sz + 4 * tj + v] = acc[u][v];
        }
        __syncthreads();
        // L[C rows, A cols] <- -C_p^-1 Tm[p]   (B_p has been consumed)
        for (int tile = t; tile < tiles; tile += NT) {
            const int pidx = tile / (tpr * tpr), rem = tile % (tpr * tpr), ti = rem / tpr, tj = rem % tpr;
            const int a0 = 2 * pidx * sz, c0 = a0 + sz;
            const float *tp = Tm + pidx * sz * sz;
            float acc[4][4] = {};
            for (int k = 0; k < sz; ++k) {
                float cv[4], tv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cv[u] = Ls[(c0 + 4 * ti + u) * TIP + c0 + k];
                    tv[u] = tp[k * sz + 4 * tj + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] += cv[u] * tv[v];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) Ls[(c0 + 4 * ti + u) * TIP + a0 + 4 * tj + v] = -acc[u][v];
        }
        __syncthreads();
    }
    // transposed store: LinvT[k * TI + r] = inverse[r][k]
    for (int idx = t; idx < TI * TI; idx += NT) {
        const int r = idx % TI, k = idx / TI;
        LinvT[idx] = Ls[r * TIP + k];
    }
}

// U12 = L11^-1 A12 for columns [c0, ncols_total), one 64-column strip per workgroup.  Wave w owns rows
// 32w .. 32w+31 of the strip (lane = column); for each k it reads those 32
// entries of column k of L11^-1 as eight 16-byte broadcast LDS reads.
__global__ __launch_bounds__(NT) void trsm_apply_kernel(float *A, long lda, int k0, int kb, int c0, int ncols_total,
                                                        const float *LinvT, const int *info)
{
    if (*info != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *Lt = reinterpret_cast<float *>(smem_raw);   // TI x TP4: Lt[k][r] = L11^-1[r][k]
    float *Bs = Lt + TI * TP4;                         // TI x 65 strip (rows >= kb zero)
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int cbase = c0 + blockIdx.x * 64;
    g2s_block<float, 16>(Lt, TP4, LinvT, TI, TI, TI, TI, TI);
    g2s_block<float, 8>(Bs, 65, A + (long)k0 * lda + cbase, lda, TI, 64, kb, ncols_total - cbase);
    __syncthreads();
    typedef float f4v __attribute__((ext_vector_type(4)));
    const int rbase = wave * 32;
    float acc[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) acc[i] = 0.0f;
#pragma unroll 2
    for (int k = 0; k < kb; ++k) {
        const float bk = Bs[k * 65 + lane];
        const f4v *lp = reinterpret_cast<const f4v *>(Lt + k * TP4 + rbase);
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const f4v l = lp[v];
            acc[4 * v] += l[0] * bk;
            acc[4 * v + 1] += l[1] * bk;
            acc[4 * v + 2] += l[2] * bk;
            acc[4 * v + 3] += l[3] * bk;
        }
    }
    const int c = cbase + lane;
    if (c < ncols_total) {
#pragma unroll
        for (int i = 0; i < 32; ++i)
            if (rbase + i < kb) A[(long)(k0 + rbase + i) * lda + c] = acc[i];
    }
}

// U12 = L11^-1 A12 for columns [c0, ncols_total) of the kb rows of the panel.
// One workgroup per 64-column strip; L11 (strict lower, packed by rows) and
// the strip live in LDS; rows are solved in blocks of 8.
template <typename T>
__global__ __launch_bounds__(NT) void trsm_llnu_kernel(T *A, long lda, int k0, int kb, int c0, int ncols_total,
                                                       const int *info)
{
    if (*info != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *Ls = reinterpret_cast<T *>(smem_raw);        // packed strict lower: row i at i*(i-1)/2
    T *Bs = Ls + (size_t)kb * (kb - 1) / 2;         // kb x 64
    const int t = threadIdx.x, lane = t & 63, rg = t >> 6;
    const int cbase = c0 + blockIdx.x * 64;
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

// Row interchanges and U12 = L11^-1 A12 in one pass, for the columns the next
// panel needs: the look-ahead's critical path between two panels.  One
// workgroup of 1024 threads per 64-column strip (lane = column):
//   1. the kb rows of the strip (before the interchanges), the strict lower
//      L11 (packed by rows) and the interchange list go to LDS;
//   2. every moved row's new value is read into registers: a row that ends in
//      the top kb rows comes from the strip or from below the panel; a row that
//      ends below comes from the strip (each interchange of partial pivoting
//      involves a top row, so a row below only ever receives a top row);
//   3. after every read has landed, the values are written: top rows into the
//      strip, rows below straight to A (so the update that follows sees them);
//   4. forward substitution with the unit-lower L11, rows in blocks of 8, the
//      rows below a block updated by all 16 waves; solved rows go to a second
//      LDS strip;
//   5. the solved strip back to rows k0 .. k0+kb-1.
constexpr int NT_LT = 1024;
template <typename T>
__global__ __launch_bounds__(NT_LT) void laswp_trsm_kernel(T *__restrict__ A, long lda, int k0, int kb, int c0, int c1,
                                                           const T *__restrict__ L, const int *pairs,
                                                           const int *npairs, const int *info)
{
    if (*info != 0) return;
    constexpr int NW = NT_LT / 64, PMAX = (2 * SCIRS_KB_MAX + NW - 1) / NW;   // interchanges per thread
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int *sp = reinterpret_cast<int *>(smem_raw);                                   // 2 * np
    T *Bs = reinterpret_cast<T *>(smem_raw + (size_t)4 * SCIRS_KB_MAX * sizeof(int));   // kb x 64
    T *Xs = Bs + (size_t)kb * 64;   // kb x 64 solved rows (separate, so no wave reads a row being replaced)
    // FP32: L11 in LDS too (pitch kb+4), read by broadcast; FP64 (no room): from memory, scalar loads
    constexpr bool LDS_L = sizeof(T) == 4;
    const int lp = kb + 4;
    T *Ls = Xs + (size_t)kb * 64;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int c = c0 + blockIdx.x * 64 + lane;
    const bool ok = c < c1;
    const int np = *npairs;
    const int top_end = k0 + kb;
    for (int q = t; q < 2 * np; q += NT_LT) sp[q] = pairs[q];
    if (LDS_L) g2s_block<T, 8>(Ls, lp, L, lda, kb, kb, kb, kb);
    {
        constexpr int U = 8;   // loads of a thread in flight before its LDS stores
        for (int r0 = wave; r0 < kb; r0 += U * NW) {
            T bv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r = r0 + u * NW;
                bv[u] = (ok && r < kb) ? A[(long)(k0 + r) * lda + c] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (r0 + u * NW < kb) Bs[(r0 + u * NW) * 64 + lane] = bv[u];
        }
    }
    __syncthreads();
    T v[PMAX];
#pragma unroll
    for (int u = 0; u < PMAX; ++u) {
        const int q = wave + u * NW;
        v[u] = (T)0;
        if (q < np) {
            const int src = sp[2 * q + 1];
            if (src < top_end) v[u] = Bs[(src - k0) * 64 + lane];
            else if (ok) v[u] = A[(long)src * lda + c];
        }
    }
    // every read of A above has returned before any thread writes a row below
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PMAX; ++u) {
        const int q = wave + u * NW;
        if (q < np) {
            const int dst = sp[2 * q];
            if (dst < top_end) Bs[(dst - k0) * 64 + lane] = v[u];
            else if (ok) A[(long)dst * lda + c] = v[u];
        }
    }
    __syncthreads();
    // L11 is read straight from memory at wave-uniform addresses (scalar loads,
    // cached): every lane of a wave needs the same L entries
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    for (int ib = 0; ib < kb; ib += 8) {
        const int ie = min(ib + 8, kb);
        // every wave solves the 8 x 8 diagonal block itself, in registers
        // (x_q -= sum_{p<q} L[ib+q][ib+p] x_p), so no barrier separates the
        // diagonal solve from the update of the rows below it
        T x[8];
        if (ie - ib == 8) {
            // full block, branch-free so every wave issues all its loads before waiting
#pragma unroll
            for (int q = 0; q < 8; ++q) x[q] = Bs[(ib + q) * 64 + lane];
            T ld[8][8];
#pragma unroll
            for (int q = 1; q < 8; ++q)
#pragma unroll
                for (int pp = 0; pp < 8; ++pp)
                    ld[q][pp] = LDS_L ? Ls[(ib + q) * lp + ib + pp] : L[(long)(ib + q) * lda + ib + pp];
#pragma unroll
            for (int q = 1; q < 8; ++q)
#pragma unroll
                for (int pp = 0; pp < q; ++pp) x[q] -= ld[q][pp] * x[pp];
            if (wv == 0) {
#pragma unroll
                for (int q = 0; q < 8; ++q) Xs[(ib + q) * 64 + lane] = x[q];
            }
            // rows below the block, four per wave at a time (row index clamped for the
            // loads, so they need no branch; rows past kb are not stored)
            for (int r0 = ie + wv; r0 < kb; r0 += 4 * NW) {
                T acc[4], lr[4][8];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int rr = min(r0 + u * NW, kb - 1);
                    acc[u] = Bs[rr * 64 + lane];
#pragma unroll
                    for (int q = 0; q < 8; ++q) lr[u][q] = LDS_L ? Ls[rr * lp + ib + q] : L[(long)rr * lda + ib + q];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int q = 0; q < 8; ++q) acc[u] -= lr[u][q] * x[q];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (r0 + u * NW < kb) Bs[(r0 + u * NW) * 64 + lane] = acc[u];
            }
        } else {
            // the last, partial block of a ragged panel
            for (int q = 0; q < ie - ib; ++q) {
                T xq = Bs[(ib + q) * 64 + lane];
                for (int pp = 0; pp < q; ++pp) xq -= L[(long)(ib + q) * lda + ib + pp] * x[pp];
                x[q] = xq;
                if (wv == 0) Xs[(ib + q) * 64 + lane] = xq;
            }
        }
        __syncthreads();
    }
    if (ok)
        for (int i = wave; i < kb; i += NW) A[(long)(k0 + i) * lda + c] = Xs[i * 64 + lane];
}

// ---- register-resident FP64 panel for short panels (m <= 512 rows, kb <= 32)
// One workgroup of 512 threads (two waves per SIMD, which hide each other's
// latencies); thread t holds the row that starts at position t in registers.
// Rows never move: a position register per row follows the interchanges
// (lapack.rs:124-131 swaps positions j and p), and the search of
// scirs2-linalg/src/lapack.rs:95-104 (largest |a_ij| over positions >= j, the
// first position on ties) is a wave arg-max plus an eight-way pick.  The
// column loop is a real loop, not unrolled: every row rotates left by one
// column per step (the update writes column c + 1 into register c, and the
// step's multiplier goes to the last register), so the pivot column is always
// register 0 and every step runs the same code from the instruction cache.
// Per column one barrier: each wave's candidate publishes its row
// (double-buffered by parity); the rows below divide by the pivot (true
// division, lapack.rs:138) and subtract their multiple (lapack.rs:140-142) —
// a row that does not take part, or a column that already holds a
// multiplier, gets a zero factor, which leaves it exactly unchanged.  The
// (destination, source) list of the moved rows comes from the final
// positions.
constexpr int GR_W = 32, GR_NT = 512, GR_MMAX = GR_NT, GR_NW = GR_NT / 64;
struct GrXch {
    double slot[2][GR_NW][GR_W];
    __attribute__((aligned(16))) double cv[2][GR_NW];
    __attribute__((aligned(16))) int ci[2][GR_NW];
    int label[GR_MMAX];
};

__global__ __launch_bounds__(GR_NT) void getf2_regs_kernel(double *A, long lda, int k0, int m, int kb, int *pairs,
                                                           int *npairs, int *info, double *info_val)
{
    if (*info != 0) return;
    __shared__ GrXch X;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    double *P = A + (long)k0 * lda + k0;
    double a[GR_W];
    int pos = t;
    {
        const double *src = P + (long)min(pos, m - 1) * lda;
#pragma unroll
        for (int c = 0; c < GR_W; ++c) a[c] = (pos < m && c < kb) ? src[c] : 0.0;
        if (pos >= m) pos = GR_MMAX;   // no row: never a candidate, never updated
    }
    bool stop = false;
#pragma unroll 1
    for (int j = 0; j < GR_W; ++j) {
        const int par = j & 1;
        const bool live = j < kb && !stop;   // uniform
        double l = 0.0;
        double pr[GR_W];
        bool below = false;
        if (live) {
            double bv = 0.0;
            int bi = -1;
            if (pos >= j && pos < m) { bv = fabs(a[0]); bi = pos; }
            // Fast wave arg-max on a 32-bit key: the top 22 bits of |v| (exponent
            // and 10 mantissa bits), a valid bit and the inverted position.  It is
            // exact unless another lane's |v| shares the winner's top 22 bits (then
            // the exact double arg-max runs instead).
            const unsigned hi = (unsigned)((unsigned long long)__double_as_longlong(bv) >> 32) & 0x7FFFFFFFu;
            const unsigned key = bi < 0 ? 0u : ((hi & ~0x3FFu) | 0x200u | (unsigned)(GR_MMAX - 1 - bi));
            unsigned km = key;
            km = max(km, (unsigned)__builtin_amdgcn_update_dpp(0, (int)km, 0x111, 0xf, 0xf, false));
            km = max(km, (unsigned)__builtin_amdgcn_update_dpp(0, (int)km, 0x112, 0xf, 0xf, false));
            km = max(km, (unsigned)__builtin_amdgcn_update_dpp(0, (int)km, 0x114, 0xf, 0xf, false));
            km = max(km, (unsigned)__builtin_amdgcn_update_dpp(0, (int)km, 0x118, 0xf, 0xf, false));
            km = max(km, (unsigned)__builtin_amdgcn_update_dpp(0, (int)km, 0x142, 0xa, 0xf, false));
            km = max(km, (unsigned)__builtin_amdgcn_update_dpp(0, (int)km, 0x143, 0xc, 0xf, false));
            km = (unsigned)__builtin_amdgcn_readlane((int)km, 63);
            const unsigned long long same = __ballot(bi >= 0 && (key & ~0x3FFu) == (km & ~0x3FFu));
            if (km != 0u && __popcll(same) == 1) {
                bi = GR_MMAX - 1 - (int)(km & 0x1FFu);
                const int src = __builtin_amdgcn_readfirstlane(__ffsll((long long)same) - 1);
                bv = readlane_val(bv, src);
            } else {
                wave_argmax(bv, bi);
            }
            if (bi >= 0 && pos == bi)
#pragma unroll
                for (int c = 0; c < GR_W; c += 2)
                    *reinterpret_cast<double2 *>(&X.slot[par][wave][c]) = double2{a[c], a[c + 1]};
            if (lane == 0) { X.cv[par][wave] = bv; X.ci[par][wave] = bi; }
            __syncthreads();
            int cis[GR_NW];
            double cvs[GR_NW];
#pragma unroll
            for (int w = 0; w < GR_NW; w += 4) {
                const int4 q = *reinterpret_cast<const int4 *>(&X.ci[par][w]);
                cis[w] = q.x; cis[w + 1] = q.y; cis[w + 2] = q.z; cis[w + 3] = q.w;
            }
#pragma unroll
            for (int w = 0; w < GR_NW; w += 2) {
                const double2 q = *reinterpret_cast<const double2 *>(&X.cv[par][w]);
                cvs[w] = q.x; cvs[w + 1] = q.y;
            }
            double v = cvs[0];
            int p = cis[0], pw = 0;
#pragma unroll
            for (int w = 1; w < GR_NW; ++w)
                if (cis[w] >= 0 && (p < 0 || cvs[w] > v || (cvs[w] == v && cis[w] < p))) {
                    v = cvs[w]; p = cis[w]; pw = w;
                }
            if (!(p >= 0 && v >= DBL_EPSILON)) {   // |pivot| < eps: singular (lapack.rs:106-120)
                if (t == 0) { *info = k0 + j + 1; *info_val = p >= 0 ? v : 0.0; }
                stop = true;
            } else {
                pos = pos == p ? j : (pos == j ? p : pos);   // the interchange of positions j and p
#pragma unroll
                for (int c = 0; c < GR_W; c += 2) {
                    const double2 q = *reinterpret_cast<const double2 *>(&X.slot[par][pw][c]);
                    // registers past the trailing columns hold multipliers: factor 0
                    pr[c] = c <= GR_W - 1 - j ? q.x : 0.0;
                    pr[c + 1] = c + 1 <= GR_W - 1 - j ? q.y : 0.0;
                }
                below = pos > j && pos < m;
                if (below) l = a[0] / pr[0];
            }
        }
        if (!live || stop) {
#pragma unroll
            for (int c = 0; c < GR_W; ++c) pr[c] = 0.0;
        }
        // rotate left by one column, updating the trailing columns of the rows below
        const double a0 = a[0];
#pragma unroll
        for (int c = 1; c < GR_W; ++c) a[c - 1] = fma(-l, pr[c], a[c]);
        a[GR_W - 1] = below ? l : a0;
    }
    // rows in natural column order again after GR_W rotations; write each to its position
    if (pos < m) {
        double *dst = P + (long)pos * lda;
#pragma unroll
        for (int c = 0; c < GR_W; ++c)
            if (c < kb) dst[c] = a[c];
        X.label[pos] = t;
    }
    __syncthreads();
    if (t < 64) {
        int np = 0;
        for (int base = 0; base < m; base += 64) {
            const int q = base + t;
            const bool moved = q < m && X.label[q] != q;
            const unsigned long long bl = __ballot(moved);
            const int at = np + __popcll(bl & ((1ull << t) - 1ull));
            if (moved) { pairs[2 * at] = k0 + q; pairs[2 * at + 1] = k0 + X.label[q]; }
            np += __popcll(bl);
        }
        if (t == 0) *npairs = np;
    }
}

} // namespace

int scirs_lu_panel_kb(int m, int kb, size_t elem_bytes)
{
    for (int w = kb; w >= 32; w = w > 64 ? 64 : w / 2) {
        const int rmax = elem_bytes == 8 ? panel::max_rows<double>(w) : panel::max_rows<float>(w);
        if (m <= rmax) return w;
        if (w == 32) break;
    }
    return kb;
}

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
    // Every panel's interchange count starts at zero: panels that never run (the
    // factorisation stopped at a singular column) must not leave the lists of an
    // earlier factorisation behind for scirs_perm_from_pairs.
    if (int rc = scirs_fill_int(w.npairs, npanels, 0, stream)) return rc;
    const bool lookahead = w.lookahead && w.side_stream && npanels > 1;
    if (lookahead) {
        // the side stream starts after everything already queued on the caller's stream
        SCIRS_HIP_CHECK(hipEventRecord(w.ev_start, stream));
        SCIRS_HIP_CHECK(hipStreamWaitEvent(w.side_stream, w.ev_start, 0));
        SCIRS_HIP_CHECK(hipEventRecord(w.ev_upd, stream));
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
        // the whole panel fits one workgroup: factor it there, with no exchange
        if (mrows <= rmax) { G = 1; R = mrows; }
        panel::PanelArgs<T> pa;
        pa.A = A; pa.lda = lda; pa.k0 = k0; pa.m = mrows; pa.kb = kbw; pa.R = R; pa.G = G;
        pa.xw = w.xw; pa.xrow = w.xrow; pa.xrj = w.xrj; pa.nrep = w.panel_nrep;
        if (int rc = scirs_next_seq(w, 512u, stream, &pa.seq_base)) return rc;
        pa.pairs = w.pairs + (size_t)p * 4 * SCIRS_KB_MAX; pa.npairs = w.npairs + p;
        pa.info = w.info; pa.info_val = w.info_val; pa.abort = w.abort; pa.spin_limit = w.spin_limit;
        pa.poll_sleep = w.poll_sleep;
        const size_t lds = panel::lds_bytes(R, kbw, sizeof(T));
        const bool w0 = w.panel_w0 && G > 1;   // the single-workgroup path is the all-waves variant
        const void *pk = w0 ? (const void *)panel::getf2_panel_kernel<T, true>
                                    : (const void *)panel::getf2_panel_kernel<T, false>;
        if (int rc = scirs_allow_lds(pk, lds)) return rc;
        if (std::is_same<T, float>::value && w.tslu_active) {
            // the mixed solve's FP32 factors: tournament pivoting (tslu.hip)
            scirs::ProfScope ps(SCIRS_PROF_PANEL, (double)mrows * kbw * kbw, 2.0 * mrows * kbw * sizeof(T), stream);
            if (int rc = scirs_tslu_panel((float *)A, lda, m, k0, kbw, p, w, stream)) return rc;
        } else {
            scirs::ProfScope ps(SCIRS_PROF_PANEL, (double)mrows * kbw * kbw, 2.0 * mrows * kbw * sizeof(T), stream);
            static const bool regs_panel = [] {
                const char *v = getenv("SCIRS_GPU_REGS_PANEL");
                return !(v && *v && atoi(v) == 0);
            }();
            if (std::is_same<T, double>::value && regs_panel && mrows <= GR_MMAX && kbw <= GR_W) {
                // a short FP64 panel: rows in registers, one barrier per column
                hipLaunchKernelGGL(getf2_regs_kernel, dim3(1), dim3(GR_NT), 0, stream, (double *)A, lda, k0, mrows,
                                   kbw, pa.pairs, pa.npairs, pa.info, pa.info_val);
                SCIRS_HIP_CHECK(hipGetLastError());
            } else if (G == 1) {
                // one workgroup: nothing to exchange, 512 threads share the rows
                constexpr int NTS = 2 * panel::NT;
                if (int rc = scirs_allow_lds((const void *)panel::getf2_panel_kernel<T, false, NTS>, lds)) return rc;
                hipLaunchKernelGGL((panel::getf2_panel_kernel<T, false, NTS>), dim3(1), dim3(NTS), lds, stream, pa);
                SCIRS_HIP_CHECK(hipGetLastError());
            } else if (lookahead) {
                // a normal launch, so the panel can share the device with the side
                // stream's update: its G <= #CUs workgroups all become resident once
                // that update's workgroups retire (they wait for nothing)
                if (w0)
                    hipLaunchKernelGGL((panel::getf2_panel_kernel<T, true>), dim3(G), dim3(panel::NT), lds, stream, pa);
                else
                    hipLaunchKernelGGL((panel::getf2_panel_kernel<T, false>), dim3(G), dim3(panel::NT), lds, stream, pa);
                SCIRS_HIP_CHECK(hipGetLastError());
            } else {
                void *args[] = {&pa};
                if (hipLaunchCooperativeKernel(pk, dim3(G), dim3(panel::NT), args, lds, stream) != hipSuccess) {
                    // cooperative launches can be refused while another process holds the
                    // device's cooperative resources: a normal launch of G <= #CUs
                    // workgroups on an idle device is co-resident in practice, and the
                    // bounded polls turn the rare case that is not into an error
                    (void)hipGetLastError();
                    if (w0)
                        hipLaunchKernelGGL((panel::getf2_panel_kernel<T, true>), dim3(G), dim3(panel::NT), lds, stream, pa);
                    else
                        hipLaunchKernelGGL((panel::getf2_panel_kernel<T, false>), dim3(G), dim3(panel::NT), lds, stream, pa);
                    SCIRS_HIP_CHECK(hipGetLastError());
                }
            }
        }
        const int nright = n - k0 - kbw;
        const int mbelow = m - k0 - kbw;
        // FP32: L11^-1 once (double-buffered, so the side stream's use of the
        // previous panel's inverse never races with this one), then multiplies
        const bool inv32 = std::is_same<T, float>::value && kbw <= TI && w.linv && nright > 0;
        float *linv = w.linv ? (float *)w.linv + (size_t)(p & 1) * TI * TI : nullptr;
        auto trtri = [&](hipStream_t s) -> int {
            if (!inv32) return SCIRS_OK;
            scirs::ProfScope ps(SCIRS_PROF_TRSM, (double)kbw * kbw * kbw / 3.0, 0.5 * kbw * kbw * sizeof(T), s);
            const size_t lds_i = ((size_t)TI * TIP + 64 * 64) * sizeof(float);
            if (int rc = scirs_allow_lds((const void *)trtri_unit_lower_kernel, lds_i)) return rc;
            hipLaunchKernelGGL(trtri_unit_lower_kernel, dim3(1), dim3(NT), lds_i, s,
                               (const float *)(A + (long)k0 * lda + k0), lda, kbw, linv, (const int *)w.info);
            SCIRS_HIP_CHECK(hipGetLastError());
            return SCIRS_OK;
        };
        // row interchanges on nv columns (the first nleft from column 0, the rest from right0)
        auto laswp = [&](int nleft, int right0, int nv, hipStream_t s) -> int {
            if (nv <= 0) return SCIRS_OK;
            const size_t lds_sw = (size_t)4 * SCIRS_KB_MAX * sizeof(int) + (size_t)2 * kbw * 64 * sizeof(T);
            if (int rc = scirs_allow_lds((const void *)laswp_kernel<T>, lds_sw)) return rc;
            scirs::ProfScope ps(SCIRS_PROF_LASWP, 0.0, 4.0 * kbw * nv * sizeof(T), s);
            hipLaunchKernelGGL(laswp_kernel<T>, dim3((nv + 63) / 64), dim3(NT), lds_sw, s,
                               A, lda, nleft, right0, nv, pa.pairs, pa.npairs, (const int *)w.info);
            SCIRS_HIP_CHECK(hipGetLastError());
            return SCIRS_OK;
        };
        // U12 = L11^-1 A12 on columns [c0, c1)
        auto trsm = [&](int c0, int c1, hipStream_t s) -> int {
            const int nc = c1 - c0;
            if (nc <= 0) return SCIRS_OK;
            scirs::ProfScope ps(SCIRS_PROF_TRSM, (double)kbw * kbw * nc, 2.0 * kbw * nc * sizeof(T), s);
            if (inv32) {
                const size_t lds_a = ((size_t)TI * TP4 + (size_t)TI * 65) * sizeof(float);
                if (int rc = scirs_allow_lds((const void *)trsm_apply_kernel, lds_a)) return rc;
                hipLaunchKernelGGL(trsm_apply_kernel, dim3((nc + 63) / 64), dim3(NT), lds_a, s,
                                   (float *)A, lda, k0, kbw, c0, c1, (const float *)linv, (const int *)w.info);
            } else {
                const size_t lds_t = ((size_t)kbw * (kbw - 1) / 2 + (size_t)kbw * 64) * sizeof(T);
                if (int rc = scirs_allow_lds((const void *)trsm_llnu_kernel<T>, lds_t)) return rc;
                hipLaunchKernelGGL(trsm_llnu_kernel<T>, dim3((nc + 63) / 64), dim3(NT), lds_t, s,
                                   A, lda, k0, kbw, c0, c1, (const int *)w.info);
            }
            SCIRS_HIP_CHECK(hipGetLastError());
            return SCIRS_OK;
        };
        // A22[:, c0:c1) -= L21 U12[:, c0:c1)
        auto update = [&](int c0, int c1, hipStream_t s) -> int {
            const int nc = c1 - c0;
            if (nc <= 0 || mbelow <= 0) return SCIRS_OK;
            scirs::ProfScope ps(SCIRS_PROF_UPDATE, 2.0 * mbelow * nc * kbw,
                                (2.0 * mbelow * nc + (double)mbelow * kbw + (double)kbw * nc) * sizeof(T), s);
            return scirs_trailing_update<T>(mbelow, nc, kbw, A + (long)(k0 + kbw) * lda + k0, lda,
                                            A + (long)k0 * lda + c0, lda, A + (long)(k0 + kbw) * lda + c0, lda, w,
                                            s);
        };
        const int cn = k0 + kbw;   // first column right of the panel
        if (!lookahead) {
            if (int rc = trtri(stream)) return rc;
            if (int rc = laswp(k0, cn, n - kbw, stream)) return rc;
            if (int rc = trsm(cn, n, stream)) return rc;
            if (int rc = update(cn, n, stream)) return rc;
            continue;
        }
        // Look-ahead.  Main stream: only what the next panel needs (its columns
        // interchanged, solved and updated).  Side stream: the interchanges of
        // every other column and the rest of the update, overlapping the next
        // panel.  The previous panel's side-stream work updates the next panel's
        // columns, so it must land first.
        const int cr = std::min(n, cn + kb);   // end of the next panel's columns
        if (p > 0) SCIRS_HIP_CHECK(hipStreamWaitEvent(stream, w.ev_upd, 0));
        SCIRS_HIP_CHECK(hipEventRecord(w.ev_trsm, stream));
        SCIRS_HIP_CHECK(hipStreamWaitEvent(w.side_stream, w.ev_trsm, 0));
        if (int rc = laswp(k0, cr, k0 + (n - cr), w.side_stream)) return rc;
        if (cr < n)
            if (int rc = trtri(w.side_stream)) return rc;
        if (int rc = trsm(cr, n, w.side_stream)) return rc;
        if (int rc = update(cr, n, w.side_stream)) return rc;
        SCIRS_HIP_CHECK(hipEventRecord(w.ev_upd, w.side_stream));
        if (cr > cn) {
            const size_t lds_lt = (size_t)4 * SCIRS_KB_MAX * sizeof(int) + (size_t)2 * kbw * 64 * sizeof(T) +
                                  (sizeof(T) == 4 ? (size_t)kbw * (kbw + 4) * sizeof(T) : 0);
            if (int rc = scirs_allow_lds((const void *)laswp_trsm_kernel<T>, lds_lt)) return rc;
            scirs::ProfScope ps(SCIRS_PROF_TRSM, (double)kbw * kbw * (cr - cn),
                                (2.0 * kbw + 4.0 * kbw) * (cr - cn) * sizeof(T), stream);
            hipLaunchKernelGGL(laswp_trsm_kernel<T>, dim3((cr - cn + 63) / 64), dim3(NT_LT), lds_lt, stream, A, lda,
                               k0, kbw, cn, cr, (const T *)(A + (long)k0 * lda + k0), pa.pairs, pa.npairs,
                               (const int *)w.info);
            SCIRS_HIP_CHECK(hipGetLastError());
        }
        if (int rc = update(cn, cr, stream)) return rc;
    }
    if (lookahead) {
        // the caller's stream sees every update
        SCIRS_HIP_CHECK(hipEventRecord(w.ev_upd, w.side_stream));
        SCIRS_HIP_CHECK(hipStreamWaitEvent(stream, w.ev_upd, 0));
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
        // a pair outside [0, m) cannot come from this factorisation: ignored
        const bool ok0 = q0 < np && (unsigned)pp[2 * q0] < (unsigned)m && (unsigned)pp[2 * q0 + 1] < (unsigned)m;
        const bool ok1 = q1 < np && (unsigned)pp[2 * q1] < (unsigned)m && (unsigned)pp[2 * q1 + 1] < (unsigned)m;
        if (ok0) v0 = perm[pp[2 * q0 + 1]];
        if (ok1) v1 = perm[pp[2 * q1 + 1]];
        __syncthreads();
        if (ok0) perm[pp[2 * q0]] = v0;
        if (ok1) perm[pp[2 * q1]] = v1;
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
