// Tall-skinny LU with tournament pivoting (CALU: Grigori, Demmel & Xiang,
// SIAM J. Matrix Anal. Appl. 32(4), 2011) for the FP32 factorisation of the
// mixed-precision solve.
//
// The partial-pivot panel (panel.hpp) needs one device-wide exchange per
// column, because the pivot of column j is the arg-max over all rows of the
// panel (scirs2-linalg/src/lapack.rs:95-104).  The mixed solve refines its
// result to FP64 and judges it by the backward error, so its FP32 factors may
// use other pivots; the reference's pivot sequence is owed only by the FP64
// `lu` / `solve` path.  Here a panel of kb <= 128 columns is factored as
// sub-panels of TW = 32 columns.  Per sub-panel:
//
//   tour_kernel  — the rows below the sub-panel's diagonal are split into
//                  blocks of NODE = 1024 rows (one 256-thread workgroup each).
//                  Every block runs Gaussian elimination with partial pivoting
//                  on its rows (registers, four rows per thread) and keeps the
//                  TW rows it picked.  The picks of up to FAN = 32 blocks
//                  (32 x 32 = 1024 rows, original values gathered from A) are
//                  eliminated again by the last of those workgroups to finish
//                  — an arrival counter, so no workgroup ever waits for one
//                  that may not be resident — and so on up the tree until one
//                  workgroup holds the TW pivot rows and, from its own
//                  elimination, their L11 \ U11.
//                  That workgroup stages the rows the pivots displace, forms
//                  U11^-1 and L11^-1, the pivot rows' U on the panel's later
//                  columns (U12 = L11^-1 A12), and composes the panel's row
//                  interchanges (the (destination, source) list of getrf).
//   apply_kernel — every row of the sub-panel, in parallel: the pivot rows get
//                  their final contents, every other row its multipliers
//                  L21 = A21 U11^-1 and the update A22 -= L21 U12 of the
//                  panel's later columns; displaced rows move to the positions
//                  the pivots left.
//
// Two launches per sub-panel replace 32 device-wide exchanges.  The pivot
// search keys are |a| with its 11 lowest bits replaced by a valid bit and the
// row index (the lowest row wins among values equal in their top 20 bits); the
// pivot is within 2^-12 relative of the column maximum, which is all the
// tournament needs.
#include "common.hpp"
#include "kernels.hpp"
#include "api_internal.hpp"
#include "prof.hpp"

#include <algorithm>
#include <type_traits>
#include <utility>

namespace tslu {

typedef unsigned long long u64;
constexpr int TW = 32;              // sub-panel width
constexpr int NT = 256;             // threads per workgroup
constexpr int RPT = 4;              // rows per thread in an elimination
constexpr int NODE = NT * RPT;      // rows per tree node
constexpr int FAN = NODE / TW;      // children per tree node
constexpr int NW = NT / 64;         // waves per workgroup
constexpr int KBT = 128;            // widest panel
constexpr int AR = 32;              // rows per apply workgroup

static_assert(NW == 4, "the key gather reads four wave keys as one uint4");
constexpr int IDXB = 10;            // bits of a node-local row index in a search key
static_assert(NODE <= (1 << IDXB), "row indexes live in the key's low bits");

__device__ __forceinline__ void st_agent(u64 *p, u64 v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 ld_agent(const u64 *p)
{
    return __hip_atomic_load(const_cast<u64 *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// search key of value x in local row r: larger |x| wins, then the lower row;
// 0 is "no candidate"
__device__ __forceinline__ unsigned mkkey(float x, int r)
{
    constexpr unsigned low = (2u << IDXB) - 1u;   // the valid bit and the row index
    return (__float_as_uint(x) & 0x7FFFFFFFu & ~low) | (1u << IDXB) | (unsigned)(NODE - 1 - r);
}

template <int CTRL, int RM>
__device__ __forceinline__ unsigned dmax(unsigned v)
{
    return max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, RM, 0xf, false));
}

// wave-wide maximum (row shifts 1, 2, 4, 8, then the row broadcasts 15 and 31:
// lane 63 ends with the maximum)
__device__ __forceinline__ unsigned wave_max_u32(unsigned v)
{
    v = dmax<0x111, 0xf>(v);
    v = dmax<0x112, 0xf>(v);
    v = dmax<0x114, 0xf>(v);
    v = dmax<0x118, 0xf>(v);
    v = dmax<0x142, 0xa>(v);
    v = dmax<0x143, 0xc>(v);
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Rows of one elimination: thread t holds local rows t + x NT (x < RPT) in registers.
struct Rows {
    float v[RPT][TW];
    int step[RPT];   // -1 not yet picked, >= 0 the step that picked it, -2 no row
};

// v <- columns c0 .. c0 + w - 1 of the given row of A (zero beyond w)
__device__ __forceinline__ void load_row_masked(float (&v)[TW], const float *src, int w)
{
#pragma unroll
    for (int c = 0; c < TW; ++c) v[c] = (c < w) ? src[c] : 0.0f;
}

// All rows of the thread: a full-width sub-panel issues all 32 16-byte loads
// before using any (one memory latency, not 32)
__device__ __forceinline__ void rows_load(Rows &R, const float *A, long lda, int c0, int w, const int *rid, int nr)
{
    const int t = threadIdx.x;
    bool have[RPT];
    const float *src[RPT];
#pragma unroll
    for (int x = 0; x < RPT; ++x) {
        const int i = t + x * NT;
        have[x] = i < nr;
        src[x] = A + (long)(have[x] ? rid[i] : rid[0]) * lda + c0;
        R.step[x] = have[x] ? -1 : -2;
    }
    if (w == TW) {
        float4 q[RPT][TW / 4];
#pragma unroll
        for (int x = 0; x < RPT; ++x)
#pragma unroll
            for (int c4 = 0; c4 < TW / 4; ++c4) q[x][c4] = reinterpret_cast<const float4 *>(src[x])[c4];
#pragma unroll
        for (int x = 0; x < RPT; ++x)
#pragma unroll
            for (int c4 = 0; c4 < TW / 4; ++c4) {
                R.v[x][4 * c4] = have[x] ? q[x][c4].x : 0.0f;
                R.v[x][4 * c4 + 1] = have[x] ? q[x][c4].y : 0.0f;
                R.v[x][4 * c4 + 2] = have[x] ? q[x][c4].z : 0.0f;
                R.v[x][4 * c4 + 3] = have[x] ? q[x][c4].w : 0.0f;
            }
    } else {
#pragma unroll
        for (int x = 0; x < RPT; ++x) {
            load_row_masked(R.v[x], src[x], w);
            if (!have[x])
#pragma unroll
                for (int c = 0; c < TW; ++c) R.v[x][c] = 0.0f;
        }
    }
}

struct Xch {
    unsigned key[2][NW];
    float4 slot[2][NW][TW / 4];
    int sel[TW];
};

// Gaussian elimination with partial pivoting on the rows in R (columns 0..w-1).
// Step j picks the unpicked row with the largest key of column j (one barrier:
// each wave publishes its best key and that row's columns j.., every thread
// reads the four keys and the winner's row) and every unpicked row subtracts
// its multiple of the pivot row from columns j+1.. .  The 64 steps are unrolled
// (column j is a register, not an index) and never branched around: a step
// with nothing to pick (past w, or no unpicked row left) runs with zero keys
// and a zero multiplier.  Picked rows stop changing.  Returns the number of
// picks; X.sel[j] is the local row picked at step j.  A picked row holds its
// multipliers (L) in columns < j and its U in columns >= j; the winner
// publishes the whole row, and with U != nullptr it is kept in U[j][0..TW-1]:
// L11 \ U11 of the picked rows in pick order.
template <int J>
__device__ __forceinline__ void gepp_step(Rows &R, int w, Xch &X, float (*U)[TW + 4], int &nsel)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int par = J & 1, J4 = J & ~3;
    const bool live = J < w;
    unsigned k[RPT];
    unsigned kk = 0u;
#pragma unroll
    for (int x = 0; x < RPT; ++x) {
        k[x] = (live && R.step[x] == -1) ? mkkey(R.v[x][J], t + x * NT) : 0u;
        kk = max(kk, k[x]);
    }
    const unsigned wm = wave_max_u32(kk);
    if (lane == 0) X.key[par][wave] = wm;
    float4 *slot = X.slot[par][wave];
#pragma unroll
    for (int x = 0; x < RPT; ++x) {
        if (wm != 0u && k[x] == wm) {
            // the root keeps the whole row (its L too); elsewhere only columns >= J are read
            if (U != nullptr) {
#pragma unroll
                for (int c4 = 0; c4 < J4 / 4; ++c4)
                    slot[c4] = make_float4(R.v[x][4 * c4], R.v[x][4 * c4 + 1], R.v[x][4 * c4 + 2], R.v[x][4 * c4 + 3]);
            }
#pragma unroll
            for (int c4 = J4 / 4; c4 < TW / 4; ++c4)
                slot[c4] = make_float4(R.v[x][4 * c4], R.v[x][4 * c4 + 1], R.v[x][4 * c4 + 2], R.v[x][4 * c4 + 3]);
        }
    }
    __syncthreads();
    const uint4 kv = *reinterpret_cast<const uint4 *>(X.key[par]);
    const unsigned gm = (unsigned)__builtin_amdgcn_readfirstlane((int)max(max(kv.x, kv.y), max(kv.z, kv.w)));
    const bool pick = gm != 0u;
    if (!pick && J < nsel) nsel = J;   // no unpicked row left (or past w)
    const int ww = __builtin_amdgcn_readfirstlane(kv.x == gm ? 0 : kv.y == gm ? 1 : kv.z == gm ? 2 : 3);
    const int rr = pick ? NODE - 1 - (int)(gm & (NODE - 1)) : -1;
    const float4 *ps = X.slot[par][ww];
    float p[TW];
#pragma unroll
    for (int c4 = J4 / 4; c4 < TW / 4; ++c4) {
        const float4 q = ps[c4];
        p[4 * c4] = q.x; p[4 * c4 + 1] = q.y; p[4 * c4 + 2] = q.z; p[4 * c4 + 3] = q.w;
    }
    if (pick && U != nullptr && t < TW) U[J][t] = reinterpret_cast<const float *>(ps)[t];
    if (pick && t == 0) X.sel[J] = rr;
#pragma unroll
    for (int x = 0; x < RPT; ++x) R.step[x] = (t + x * NT == rr) ? J : R.step[x];
    const float rp = (pick && p[J] != 0.0f) ? __builtin_amdgcn_rcpf(p[J]) : 0.0f;
#pragma unroll
    for (int x = 0; x < RPT; ++x) {
        const bool un = R.step[x] == -1;
        const float l = un ? R.v[x][J] * rp : 0.0f;
        R.v[x][J] = un ? l : R.v[x][J];
#pragma unroll
        for (int c = J + 1; c < TW; ++c) R.v[x][c] = fmaf(-l, p[c], R.v[x][c]);
    }
}

template <int J>
__device__ __forceinline__ void gepp_steps(Rows &R, int w, Xch &X, float (*U)[TW + 4], int &nsel)
{
    if constexpr (J < TW) {
        gepp_step<J>(R, w, X, U, nsel);
        gepp_steps<J + 1>(R, w, X, U, nsel);
    }
}

__device__ __forceinline__ int rows_gepp(Rows &R, int w, Xch &X, float (*U)[TW + 4])
{
    int nsel = w;
    gepp_steps<0>(R, w, X, U, nsel);
    return nsel;
}

// entry (i, j) of the upper-triangular matrix held in M (pitch ldm), or of the
// transpose of the lower-triangular one (TRANS)
template <bool TRANS, bool UNIT>
__device__ __forceinline__ float tri_u(const float *M, int ldm, int i, int j)
{
    return TRANS ? M[j * ldm + i] : M[i * ldm + j];
}

// One wave: X = U^-1, lane c computing column c by right-looking substitution
// (once x_i is final its multiples of column i of U leave the rows above it:
// independent updates, and a reciprocal instead of a division for the
// diagonal — these factors only precondition refinement).
template <bool TRANS, bool UNIT>
__device__ __forceinline__ void tri_inv_wave(const float *M, int ldm, float *X, int ldx)
{
    const int c = threadIdx.x & 63;
    if (c >= TW) return;
    float x[TW];
#pragma unroll
    for (int i = 0; i < TW; ++i) x[i] = (i == c) ? 1.0f : 0.0f;
#pragma unroll
    for (int i = TW - 1; i >= 0; --i) {
        if (!UNIT) x[i] *= __builtin_amdgcn_rcpf(tri_u<TRANS, UNIT>(M, ldm, i, i));
#pragma unroll
        for (int k = 0; k < i; ++k) x[k] = fmaf(-tri_u<TRANS, UNIT>(M, ldm, k, i), x[i], x[k]);
    }
#pragma unroll
    for (int i = 0; i < TW; ++i) X[i * ldx + c] = x[i];
}

struct TourArgs {
    float *A; long lda;
    int k0, kb;         // the panel: columns (and first row position) k0, width kb
    int c0, w;          // this sub-panel: columns c0 .. c0 + w - 1, rows c0 .. n - 1
    int n;
    int nlev;           // tree levels above the leaves (0: one leaf, the root)
    int lev_n[TSLU_MAXLEV + 1];
    u64 *cand; int *cnt; unsigned seq;
    float *uinv, *urest, *top, *stage;
    int *vlist, *map;
    int first, last;
    int *pairs, *npairs;
    int *info; double *info_val; int *abort;
    unsigned spin_limit;
};

constexpr int LP = TW + 4;     // LDS pitch of the sub-panel-wide blocks (16-byte rows)
constexpr int PP = KBT + 4;    // LDS pitch of the pivot rows' panel rows

// LDS of the root's work (dynamic; the tree's workgroups carry it too)
struct RootLds {
    float lu[TW][LP];       // L11 \ U11 of the pivot rows in pick order, from the root's elimination
    float uinv[TW][LP];     // U11^-1
    float linvT[TW][LP];    // (L11^-1)^T
    float pr[TW][PP];       // the pivot rows' panel rows (earlier sub-panels' L, then U12 on the later columns)
    int gp[TW];
    int mark[TW];
    int dlist[TW], vlist[TW];
    int mpos[TSLU_MAP], mlab[TSLU_MAP];
    int mv_dst[2 * TW], mv_src[2 * TW], mv_lab[2 * TW];
    int nd, nmv, mcount, bad, npairs;
};

__device__ __forceinline__ void root_work(const TourArgs &a, int nsel, const int *rid, Xch &X, RootLds &L)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int w = a.w, r0 = a.c0, cur0 = a.c0 - a.k0, rest0 = cur0 + w, nrest = a.kb - rest0, kb = a.kb;
    // the earlier sub-panels' interchange map (final when this launch began),
    // read now so that its latency passes under steps 1-5 (barriers here wait
    // for LDS only, not for global loads)
    static_assert(TSLU_MAP == 2 * NT, "two map entries per thread");
    int map_pos[2], map_lab[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        map_pos[x] = a.first ? 0 : a.map[t + x * NT];
        map_lab[x] = a.first ? 0 : a.map[TSLU_MAP + t + x * NT];
    }
    const int map_count = a.first ? 0 : a.map[2 * TSLU_MAP];
    // 1. pivot positions in pick order; the breakdown test |u_jj| < eps (the
    //    absolute test of lapack.rs:106-120: the mixed solve then falls back to
    //    FP64 LU, which decides singularity)
    if (t < nsel) L.gp[t] = rid[X.sel[t]];
    if (t == 0) L.bad = nsel < w ? nsel : TW;
    __syncthreads();
    if (t < nsel && t < w && !(fabsf(L.lu[t][t]) >= FLT_EPSILON)) atomicMin(&L.bad, t);
    __syncthreads();
    if (L.bad < w) {
        if (t == 0) {
            *a.info = a.c0 + L.bad + 1;
            *a.info_val = L.bad < nsel ? (double)fabsf(L.lu[L.bad][L.bad]) : 0.0;
        }
        return;
    }
    if (w < TW)   // a ragged last sub-panel: the identity beyond w
        for (int q = t; q < TW * TW; q += NT) {
            const int i = q / TW, c = q % TW;
            if (i >= w || c >= w) L.lu[i][c] = (i == c) ? 1.0f : 0.0f;
        }
    // 2. positions: top r0 .. r0 + w - 1 receive the pivots in pick order; the
    //    top rows that are not pivots (dlist) move to the positions of the
    //    pivots from below (vlist)
    if (t < TW) L.mark[t] = 0;
    __syncthreads();
    if (t < w && L.gp[t] < r0 + w) L.mark[L.gp[t] - r0] = 1;
    __syncthreads();
    if (wave == 0) {
        const bool top_is_pivot = lane >= w || L.mark[lane] != 0;
        const unsigned long long bd = __ballot(lane < w && !top_is_pivot);
        if (lane < w && !top_is_pivot) L.dlist[__popcll(bd & ((1ull << lane) - 1ull))] = r0 + lane;
        const bool below = lane < w && L.gp[lane] >= r0 + w;
        const unsigned long long bv = __ballot(below);
        if (below) L.vlist[__popcll(bv & ((1ull << lane) - 1ull))] = L.gp[lane];
        if (lane == 0) L.nd = __popcll(bd);
    }
    __syncthreads();
    const int nd = L.nd;
    // 3 + 4 side by side.  Waves 2-3: the pivot rows' panel rows into LDS and
    // the displaced rows staged whole (16-byte loads, all issued before any is
    // stored).  Waves 0-1: U11^-1 and (L11^-1)^T = the inverse of L11^T (upper,
    // unit diagonal), which need only the elimination's L11 \ U11.
    if (wave >= 2) {
        constexpr int NLD = NT / 2;                        // threads of waves 2-3
        constexpr int JOBS = 2 * TW * (KBT / 4) / NLD;     // (row, 16-byte chunk) pairs per thread
        const int tt = t - NLD;
        float4 q[JOBS];
#pragma unroll
        for (int u = 0; u < JOBS; ++u) {
            const int job = tt + u * NLD, r = job / (KBT / 4), c4 = job % (KBT / 4);
            const bool piv = r < TW;
            const int ri = piv ? r : r - TW;
            const bool ok = (piv ? ri < w : ri < nd) && 4 * c4 < kb;
            const int row = ok ? (piv ? L.gp[ri] : L.dlist[ri]) : r0;
            const float *src = a.A + (long)row * a.lda + a.k0 + 4 * c4;
            q[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (ok) {
                if (4 * c4 + 3 < kb) {
                    q[u] = *reinterpret_cast<const float4 *>(src);
                } else {
                    q[u].x = src[0];
                    if (4 * c4 + 1 < kb) q[u].y = src[1];
                    if (4 * c4 + 2 < kb) q[u].z = src[2];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < JOBS; ++u) {
            const int job = tt + u * NLD, r = job / (KBT / 4), c4 = job % (KBT / 4);
            const bool piv = r < TW;
            const int ri = piv ? r : r - TW;
            if ((piv ? ri < w : ri < nd) && 4 * c4 < kb) {
                if (piv) *reinterpret_cast<float4 *>(&L.pr[ri][4 * c4]) = q[u];
                else *reinterpret_cast<float4 *>(a.stage + ri * SCIRS_KB_MAX + 4 * c4) = q[u];
            }
        }
    } else if (wave == 0) {
        tri_inv_wave<false, false>(&L.lu[0][0], LP, &L.uinv[0][0], LP);
        if (t < nd) a.vlist[t] = L.vlist[t];
        if (t == 0) a.vlist[TW] = nd;
    } else {
        tri_inv_wave<true, true>(&L.lu[0][0], LP, &L.linvT[0][0], LP);
    }
    __syncthreads();
    // 5. U12 = L11^-1 A12 on the later columns (4 x 4 register tiles)
    constexpr int TJ = KBT / 4;
    static_assert((TW / 4) * TJ <= NT, "one U12 tile per thread");
    const int ti = t / TJ, tj = t % TJ;   // rows 4 ti .., columns 4 tj ..
    const bool have12 = ti < TW / 4 && 4 * tj < nrest && 4 * ti < w;
    float u12[4][4] = {};
    if (have12) {
#pragma unroll
        for (int k = 0; k < TW; ++k) {
            const float4 lv = *reinterpret_cast<const float4 *>(&L.linvT[k][4 * ti]);
            const float4 av = *reinterpret_cast<const float4 *>(&L.pr[k][rest0 + 4 * tj]);
            const float l4[4] = {lv.x, lv.y, lv.z, lv.w}, a4[4] = {av.x, av.y, av.z, av.w};
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) u12[u][v] = fmaf(l4[u], a4[v], u12[u][v]);
        }
    }
    __syncthreads();
    if (have12) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = 4 * ti + u;
            if (i >= w) continue;
            if (4 * tj + 3 < nrest)
                *reinterpret_cast<float4 *>(a.urest + i * KBT + 4 * tj) =
                    make_float4(u12[u][0], u12[u][1], u12[u][2], u12[u][3]);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = 4 * tj + v;
                if (c < nrest) {
                    L.pr[i][rest0 + c] = u12[u][v];
                    if (4 * tj + 3 >= nrest) a.urest[i * KBT + c] = u12[u][v];
                }
            }
        }
    }
    // the pivot rows' current columns: L11 below the diagonal, U11 on and above
    for (int q = t; q < w * w; q += NT) {
        const int i = q / w, c = q - i * w;
        L.pr[i][cur0 + c] = L.lu[i][c];
    }
    for (int q = t; q < TW * TW / 4; q += NT) {
        const int i = q / (TW / 4), c4 = q % (TW / 4);
        *reinterpret_cast<float4 *>(a.uinv + i * TW + 4 * c4) = *reinterpret_cast<const float4 *>(&L.uinv[i][4 * c4]);
    }
    __syncthreads();
    for (int q = t; q < w * (KBT / 4); q += NT) {
        const int j = q / (KBT / 4), c4 = q % (KBT / 4);
        if (4 * c4 < kb)
            *reinterpret_cast<float4 *>(a.top + j * SCIRS_KB_MAX + 4 * c4) =
                *reinterpret_cast<const float4 *>(&L.pr[j][4 * c4]);
    }
    // 6. the panel's interchanges: (position, original row) for every moved row
    if (t == 0) {
        L.mcount = map_count;
        L.nmv = 0;
    }
    __syncthreads();
    const int mc0 = map_count;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        if (t + x * NT < mc0) {
            L.mpos[t + x * NT] = map_pos[x];
            L.mlab[t + x * NT] = map_lab[x];
        }
    }
    if (t < w && L.gp[t] != r0 + t) {
        const int m = atomicAdd(&L.nmv, 1);
        L.mv_dst[m] = r0 + t;
        L.mv_src[m] = L.gp[t];
    } else if (t >= TW && t < TW + nd) {
        const int m = atomicAdd(&L.nmv, 1);
        L.mv_dst[m] = L.vlist[t - TW];
        L.mv_src[m] = L.dlist[t - TW];
    }
    __syncthreads();
    const int nmv = L.nmv;
    // each wave looks its moves up 64 map entries at a time: the original row
    // of every source, and the entry (if any) of every destination
    for (int m = wave; m < nmv; m += NW) {
        const int src = L.mv_src[m], dst = L.mv_dst[m];
        int fs = -1, fd = -1;
        for (int base = 0; base < mc0; base += 64) {
            const int e = base + lane;
            const int pos = e < mc0 ? L.mpos[e] : -1;
            const unsigned long long bs = __ballot(pos == src), bd2 = __ballot(pos == dst);
            if (fs < 0 && bs) fs = base + __ffsll(bs) - 1;
            if (fd < 0 && bd2) fd = base + __ffsll(bd2) - 1;
        }
        if (lane == 0) {
            L.mv_lab[m] = fs >= 0 ? L.mlab[fs] : src;
            L.mv_src[m] = fd;   // reused: the destination's entry, -1 when new
        }
    }
    __syncthreads();
    if (t < nmv) {
        int at = L.mv_src[t];
        if (at < 0) {
            at = atomicAdd(&L.mcount, 1);
            L.mpos[at] = L.mv_dst[t];
        }
        L.mlab[at] = L.mv_lab[t];
    }
    __syncthreads();
    const int mc = L.mcount;
    if (a.last) {
        if (t == 0) L.npairs = 0;
        __syncthreads();
        for (int e = t; e < mc; e += NT) {
            if (L.mpos[e] != L.mlab[e]) {
                const int q = atomicAdd(&L.npairs, 1);
                a.pairs[2 * q] = L.mpos[e];
                a.pairs[2 * q + 1] = L.mlab[e];
            }
        }
        __syncthreads();
        if (t == 0) *a.npairs = L.npairs;
    } else {
        for (int e = t; e < mc; e += NT) {
            a.map[e] = L.mpos[e];
            a.map[TSLU_MAP + e] = L.mlab[e];
        }
        if (t == 0) a.map[2 * TSLU_MAP] = mc;
    }
}

__global__ __launch_bounds__(NT) void tour_kernel(TourArgs a)
{
    if (*a.info != 0) return;
    __shared__ Xch X;
    __shared__ int rid[NODE];
    __shared__ int s_cnt[FAN], s_off[FAN + 1];
    __shared__ int s_go;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    RootLds &L = *reinterpret_cast<RootLds *>(smem_raw);
    const int t = threadIdx.x;
    int node = blockIdx.x, lev = 0;
    const int base = a.c0 + node * NODE;
    int nr = min(NODE, a.n - base);
    for (int i = t; i < nr; i += NT) rid[i] = base + i;
    bool eliminate = true;
    int nsel = 0;
    while (true) {
        if (eliminate) {
            __syncthreads();
            const bool root = lev == a.nlev;
            if (root)
                for (int q = t; q < TW * LP; q += NT) (&L.lu[0][0])[q] = 0.0f;
            Rows R;
            rows_load(R, a.A, a.lda, a.c0, a.w, rid, nr);
            nsel = rows_gepp(R, a.w, X, root ? L.lu : nullptr);
        }
        if (lev == a.nlev) break;
        __syncthreads();
        const int parent = node / FAN;
        const int nch = min(FAN, a.lev_n[lev] - parent * FAN);
        if (nch == 1) {
            // a sole child's picks stand for its parent's: keep them, in pick order
            const int r = t < nsel ? rid[X.sel[t]] : 0;
            __syncthreads();
            if (t < nsel) {
                rid[t] = r;
                X.sel[t] = t;
            }
            nr = nsel;
            node = parent;
            ++lev;
            eliminate = lev == a.nlev;   // the root always eliminates (it keeps U)
            continue;
        }
        // publish this node's picks (tagged words), then count the arrival
        u64 *slot = a.cand + ((size_t)lev * TSLU_MAXNODE + node) * (TW + 1);
        if (t <= nsel) {
            const unsigned val = t == 0 ? (unsigned)nsel : (unsigned)rid[X.sel[t - 1]];
            st_agent(slot + t, ((u64)a.seq << 32) | val);
        }
        __syncthreads();
        if (t == 0) {
            int *cnt = a.cnt + (lev + 1) * TSLU_MAXNODE + parent;
            const int old = __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_go = old == nch - 1;
            if (s_go) __hip_atomic_store(cnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        if (!s_go) return;   // a sibling still works: the last to arrive merges
        // the last arrival merges the children's picks: their counts, then rows
        const u64 *cb = a.cand + ((size_t)lev * TSLU_MAXNODE + parent * FAN) * (TW + 1);
        node = parent;
        ++lev;
        bool late = false;
        if (t < nch) {
            unsigned spins = 0;
            u64 wv;
            while (((wv = ld_agent(cb + t * (TW + 1))) >> 32) != a.seq)
                if (++spins > a.spin_limit) { late = true; break; }
            s_cnt[t] = late ? 0 : (int)(unsigned)wv;
        }
        if (__syncthreads_or(late)) {
            if (t == 0) {
                __hip_atomic_store(a.abort, a.c0 + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(a.info, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            return;
        }
        if (t == 0) {
            int o = 0;
            for (int c = 0; c < nch; ++c) { s_off[c] = o; o += min(s_cnt[c], TW); }
            s_off[nch] = o;
        }
        __syncthreads();
        nr = s_off[nch];
        // every (child, pick) word at once: one round trip, not one per child
        for (int q = t; q < nch * TW; q += NT) {
            const int c = q / TW, j = q - c * TW;
            if (j < s_off[c + 1] - s_off[c]) {
                unsigned spins = 0;
                u64 wv;
                while (((wv = ld_agent(cb + c * (TW + 1) + 1 + j)) >> 32) != a.seq)
                    if (++spins > a.spin_limit) { late = true; break; }
                rid[s_off[c] + j] = (int)(unsigned)wv;
            }
        }
        if (__syncthreads_or(late)) {
            if (t == 0) {
                __hip_atomic_store(a.abort, a.c0 + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(a.info, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            return;
        }
        eliminate = true;
    }
    __syncthreads();
    root_work(a, nsel, rid, X, L);
}

struct ApplyArgs {
    float *A; long lda;
    int k0, kb, c0, w, n;
    const float *uinv, *urest, *top, *stage;
    const int *vlist;
    const int *info;
};

// Rows c0 .. n-1 of the panel, AR per workgroup: the sub-panel's pivot rows
// take their final contents; every other row (in place, or a displaced row
// from the staging buffer at the position a pivot left) gets
//   L21 = A21 U11^-1   (the sub-panel's columns)
//   A22 -= L21 U12     (the panel's later columns).
// Thread (row rho, group g) computes 8 consecutive columns of each product;
// rows move as 16-byte chunks, all loads issued before any is stored.
__global__ __launch_bounds__(NT) void apply_kernel(ApplyArgs a)
{
    if (*a.info != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float(*Ui)[LP] = reinterpret_cast<float(*)[LP]>(smem_raw);   // TW x LP
    float(*Ur)[PP] = reinterpret_cast<float(*)[PP]>(Ui + TW);    // TW x PP (the later columns)
    float(*Ar)[PP] = Ur + TW;                                    // AR x PP
    float(*Lr)[LP] = reinterpret_cast<float(*)[LP]>(Ar + AR);    // AR x LP
    __shared__ int sidx[AR];
    const int t = threadIdx.x;
    const int w = a.w, cur0 = a.c0 - a.k0, rest0 = cur0 + w, nrest = a.kb - rest0, kb = a.kb;
    const int d0 = a.c0 + blockIdx.x * AR, top_end = a.c0 + w;
    for (int q = t; q < TW * TW / 4; q += NT) {
        const int r = q / (TW / 4), c4 = q % (TW / 4);
        float4 u = *reinterpret_cast<const float4 *>(a.uinv + r * TW + 4 * c4);
        float *uu = &u.x;
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (r >= w || 4 * c4 + e >= w) uu[e] = 0.0f;
        *reinterpret_cast<float4 *>(&Ui[r][4 * c4]) = u;
    }
    for (int q = t; q < TW * KBT / 4; q += NT) {
        const int r = q / (KBT / 4), c4 = q % (KBT / 4);
        float4 v = *reinterpret_cast<const float4 *>(a.urest + r * KBT + 4 * c4);
        float *vv = &v.x;
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (r >= w || 4 * c4 + e >= nrest) vv[e] = 0.0f;
        *reinterpret_cast<float4 *>(&Ur[r][4 * c4]) = v;
    }
    if (t < AR) sidx[t] = -1;
    __syncthreads();
    const int nv = a.vlist[TW];
    for (int i = t; i < nv; i += NT) {
        const int pos = a.vlist[i];
        if (pos >= d0 && pos < d0 + AR) sidx[pos - d0] = i;
    }
    __syncthreads();
    // rows in: the top rows' final contents, staged rows, or the row in place
    // (from the sub-panel's first column on)
    constexpr int JOBS = AR * (KBT / 4) / NT;
    float4 q[JOBS];
#pragma unroll
    for (int u = 0; u < JOBS; ++u) {
        const int job = t + u * NT, rho = job / (KBT / 4), c4 = job % (KBT / 4), d = d0 + rho;
        q[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (d < a.n && 4 * c4 < kb) {
            const float *src = nullptr;
            if (d < top_end) src = a.top + (d - a.c0) * SCIRS_KB_MAX;
            else if (sidx[rho] >= 0) src = a.stage + sidx[rho] * SCIRS_KB_MAX;
            else if (4 * c4 >= cur0) src = a.A + (long)d * a.lda + a.k0;
            if (src) {
                src += 4 * c4;
                if (4 * c4 + 3 < kb) {
                    q[u] = *reinterpret_cast<const float4 *>(src);
                } else {
                    q[u].x = src[0];
                    if (4 * c4 + 1 < kb) q[u].y = src[1];
                    if (4 * c4 + 2 < kb) q[u].z = src[2];
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < JOBS; ++u) {
        const int job = t + u * NT, rho = job / (KBT / 4), c4 = job % (KBT / 4);
        *reinterpret_cast<float4 *>(&Ar[rho][4 * c4]) = q[u];
    }
    __syncthreads();
    // thread (row rho, group g): CL columns of L21 and CR of the later columns
    constexpr int CL = TW / 8, CR = (KBT - TW) / 8;
    static_assert(CL % 4 == 0 && CR % 4 == 0, "16-byte LDS reads");
    const int rho = t >> 3, g = t & 7, d = d0 + rho;
    const bool work = d >= top_end && d < a.n;
    if (work) {
        float l[CL] = {};
#pragma unroll
        for (int k = 0; k < TW; ++k) {
            const float av = Ar[rho][cur0 + k];
#pragma unroll
            for (int e4 = 0; e4 < CL / 4; ++e4) {
                const float4 u = *reinterpret_cast<const float4 *>(&Ui[k][CL * g + 4 * e4]);
                l[4 * e4] = fmaf(av, u.x, l[4 * e4]);
                l[4 * e4 + 1] = fmaf(av, u.y, l[4 * e4 + 1]);
                l[4 * e4 + 2] = fmaf(av, u.z, l[4 * e4 + 2]);
                l[4 * e4 + 3] = fmaf(av, u.w, l[4 * e4 + 3]);
            }
        }
#pragma unroll
        for (int e4 = 0; e4 < CL / 4; ++e4)
            *reinterpret_cast<float4 *>(&Lr[rho][CL * g + 4 * e4]) =
                make_float4(l[4 * e4], l[4 * e4 + 1], l[4 * e4 + 2], l[4 * e4 + 3]);
    }
    __syncthreads();
    if (work) {
        if (CR * g < nrest) {
            float r[CR];
#pragma unroll
            for (int e = 0; e < CR; ++e) r[e] = Ar[rho][rest0 + CR * g + e];
#pragma unroll
            for (int k = 0; k < TW; ++k) {
                const float lv = Lr[rho][k];
#pragma unroll
                for (int e4 = 0; e4 < CR / 4; ++e4) {
                    const float4 u = *reinterpret_cast<const float4 *>(&Ur[k][CR * g + 4 * e4]);
                    r[4 * e4] = fmaf(-lv, u.x, r[4 * e4]);
                    r[4 * e4 + 1] = fmaf(-lv, u.y, r[4 * e4 + 1]);
                    r[4 * e4 + 2] = fmaf(-lv, u.z, r[4 * e4 + 2]);
                    r[4 * e4 + 3] = fmaf(-lv, u.w, r[4 * e4 + 3]);
                }
            }
#pragma unroll
            for (int e = 0; e < CR; ++e) Ar[rho][rest0 + CR * g + e] = r[e];
        }
#pragma unroll
        for (int e = 0; e < CL; ++e) Ar[rho][cur0 + CL * g + e] = Lr[rho][CL * g + e];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < JOBS; ++u) {
        const int job = t + u * NT, rho2 = job / (KBT / 4), c4 = job % (KBT / 4), dd = d0 + rho2;
        if (dd >= a.n || 4 * c4 >= kb) continue;
        if (!(4 * c4 >= cur0 || dd < top_end || sidx[rho2] >= 0)) continue;
        float *dst = a.A + (long)dd * a.lda + a.k0 + 4 * c4;
        const float4 v = *reinterpret_cast<const float4 *>(&Ar[rho2][4 * c4]);
        if (4 * c4 + 3 < kb) {
            *reinterpret_cast<float4 *>(dst) = v;
        } else {
            dst[0] = v.x;
            if (4 * c4 + 1 < kb) dst[1] = v.y;
            if (4 * c4 + 2 < kb) dst[2] = v.z;
        }
    }
}

} // namespace tslu

int scirs_tslu_panel(float *A, long lda, int n, int k0, int kb, int panel_index, ScirsWork &w, hipStream_t s)
{
    using namespace tslu;
    if (kb <= 0 || kb > KBT || (lda & 3) || (k0 & 3)) {
        scirs_set_error("tslu panel: unsupported shape (kb=%d lda=%ld k0=%d)", kb, lda, k0);
        return SCIRS_ERR_ARG;
    }
    const size_t lds_root = sizeof(RootLds);
    if (int rc = scirs_allow_lds((const void *)tour_kernel, lds_root)) return rc;
    const size_t lds_apply = ((size_t)TW * LP + (size_t)TW * PP + (size_t)AR * PP + (size_t)AR * LP) * sizeof(float);
    if (int rc = scirs_allow_lds((const void *)apply_kernel, lds_apply)) return rc;
    const int nsub = (kb + TW - 1) / TW;
    for (int sp = 0; sp < nsub; ++sp) {
        const int c0 = k0 + sp * TW;
        const int wc = std::min(TW, k0 + kb - c0);
        const int mrows = n - c0;
        TourArgs ta{};
        ta.A = A; ta.lda = lda; ta.k0 = k0; ta.kb = kb; ta.c0 = c0; ta.w = wc; ta.n = n;
        ta.lev_n[0] = (mrows + NODE - 1) / NODE;
        int nlev = 0;
        while (ta.lev_n[nlev] > 1) {
            if (nlev + 1 >= TSLU_MAXLEV || ta.lev_n[nlev] > TSLU_MAXNODE) {
                scirs_set_error("tslu panel: %d rows exceed the tournament tree", mrows);
                return SCIRS_ERR_UNSUPPORTED;
            }
            ta.lev_n[nlev + 1] = (ta.lev_n[nlev] + FAN - 1) / FAN;
            ++nlev;
        }
        ta.nlev = nlev;
        ta.cand = w.tp_cand; ta.cnt = w.tp_cnt;
        if (int rc = scirs_next_seq(w, 1u, s, &ta.seq)) return rc;
        ta.uinv = w.tp_uinv; ta.urest = w.tp_urest; ta.top = w.tp_top; ta.stage = w.tp_stage;
        ta.vlist = w.tp_vlist; ta.map = w.tp_map;
        ta.first = sp == 0; ta.last = sp == nsub - 1;
        ta.pairs = w.pairs + (size_t)panel_index * 4 * SCIRS_KB_MAX;
        ta.npairs = w.npairs + panel_index;
        ta.info = w.info; ta.info_val = w.info_val; ta.abort = w.abort; ta.spin_limit = w.spin_limit;
        hipLaunchKernelGGL(tour_kernel, dim3(ta.lev_n[0]), dim3(NT), lds_root, s, ta);
        SCIRS_HIP_CHECK(hipGetLastError());
        ApplyArgs aa{};
        aa.A = A; aa.lda = lda; aa.k0 = k0; aa.kb = kb; aa.c0 = c0; aa.w = wc; aa.n = n;
        aa.uinv = w.tp_uinv; aa.urest = w.tp_urest; aa.top = w.tp_top; aa.stage = w.tp_stage;
        aa.vlist = w.tp_vlist; aa.info = w.info;
        hipLaunchKernelGGL(apply_kernel, dim3((mrows + AR - 1) / AR), dim3(NT), lds_apply, s, aa);
        SCIRS_HIP_CHECK(hipGetLastError());
    }
    return SCIRS_OK;
}

namespace {
int lowp_of_mode(int mode)
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

extern "C" int scirs_gpu_sgetrf_tp_dev(int n, float *dA, int64_t lda, int64_t *dperm, int mode, void *stream)
{
    using namespace scirs;
    const int lowp = lowp_of_mode(mode == SCIRS_MODE_AUTO ? SCIRS_MODE_BF16X6 : mode);
    if (n <= 0 || !dA || !dperm || lda < n || (lda & 3) || lowp < 0 || lowp == SCIRS_LOWP_FP16 ||
        lowp == SCIRS_LOWP_FP16X3) {
        scirs_set_error("sgetrf_tp_dev: invalid arguments (n=%d lda=%lld mode=%d; lda a multiple of 4, FP16 "
                        "needs the mixed solve's scaling)", n, (long long)lda, mode);
        return SCIRS_ERR_ARG;
    }
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    Context &c = ctx();
    if (int rc = c.init(0)) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (int rc = c.reset_flags(s)) return rc;
    const int saved = c.work.lowp_mode;
    c.work.lowp_mode = lowp;
    c.work.tslu_active = 1;
    int rc = scirs_getrf_device<float>(n, n, dA, lda, c.kb_f32, c.work, s);
    c.work.tslu_active = 0;
    c.work.lowp_mode = saved;
    if (rc) return rc;
    if ((rc = scirs_perm_from_pairs(n, n, c.kb_f32, c.work, dperm, s))) return rc;
    int st = 0;
    if ((rc = c.read_flags(s, &st, nullptr))) return rc;
    return st;
}
