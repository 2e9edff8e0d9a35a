// Chained triangular solve: X <- T^-1 X for T the unit-lower or upper triangle
// of a row-major matrix, in ONE launch.
//
// Same arithmetic as the substitution loops of scirs2-linalg/src/solve.rs:159-199
// (true division by the diagonal in FP64; multiplication by the reciprocal in
// FP32), with the k-sum split into 64-column blocks.  The matrix is cut into 64-row block rows; block row p is solved by one
// workgroup once the x values of every block it depends on are known.
//
//   * Workgroups take block rows in dependency order from an atomic ticket, so
//     every block a workgroup waits for is held by a workgroup that is already
//     running (or done): no deadlock whatever the grid size or dispatch order.
//   * A finished block publishes its x values as sequence-tagged 64-bit words
//     (relaxed agent-scope atomics; a value is current when its word carries
//     this launch's number), which later block rows poll: no fences, no flags,
//     one memory round trip per hand-over.  Polls are bounded and raise *abort.
//   * Waves 1-3 stream the off-diagonal tiles of their block row from HBM as
//     the x blocks arrive (each lane: 4 consecutive columns of 16 rows, so a
//     load instruction covers whole 256-byte row segments) and keep per-lane
//     partial sums.  Wave 0 stages the diagonal block and the LAST off-diagonal
//     tile in LDS, and after a single barrier does the only work on the
//     critical path: that last tile's product (one row per lane, x broadcast by
//     readlane) and the 64-step substitution.
//
// Algorithmic traffic per solve: the triangle once (n^2/2 elements) plus the
// right-hand sides; the chain is bounded by ~n/64 hand-overs.
#include "api_internal.hpp"
#include "panel.hpp"
#include "prof.hpp"

namespace {

using panel::Tag;
using panel::u64;
using panel::ld_agent;
using panel::st_agent;

constexpr int TB = 64;        // rows per block row
constexpr int NTC = 256;      // 4 waves
constexpr int LDP = TB + 1;   // LDS pitch of staged tiles

template <typename T>
__device__ __forceinline__ T readlane(T v, int k);
template <>
__device__ __forceinline__ float readlane<float>(float v, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
template <>
__device__ __forceinline__ double readlane<double>(double v, int k)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Poll until every lane holds its first nvalid of the NV tagged values at src
// (NV * W consecutive words; the values past nvalid are not read and come back
// as zero, and a lane with nvalid = 0 takes no part).  PD reads of the words
// are kept in flight, issued one after another and each checked as it
// returns, so a value is seen a fraction of a memory round trip after it lands
// instead of up to a whole one.  Gives up (returns false) after spin_limit
// rounds of PD reads or when another workgroup raised *abort.
template <typename T, int NV, int PD>
__device__ bool poll_vals(const u64 *src, int nvalid, unsigned seq, T (&v)[NV], unsigned spin_limit, int *abort,
                          int *info)
{
    constexpr int W = Tag<T>::W, NWD = NV * W;
    const bool active = nvalid > 0;
    const int nw = nvalid * W;
    const unsigned long long am = __ballot(active);
    if (am == 0ull) {
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = (T)0;
        return true;
    }
    // words that are not polled are read from a live address (this lane's
    // first word, or the first active lane's) and ignored
    const int first = __ffsll(am) - 1;
    const uintptr_t sa = (uintptr_t)src;
    const uintptr_t s0 = (uintptr_t)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(sa & 0xffffffffu), first) |
                         ((uintptr_t)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(sa >> 32), first) << 32);
    const u64 *p = active ? src : reinterpret_cast<const u64 *>(s0);
    u64 w[PD][NWD];
#pragma unroll
    for (int d = 0; d < PD; ++d)
#pragma unroll
        for (int i = 0; i < NWD; ++i) w[d][i] = ld_agent(p + (i < nw ? i : 0));
    unsigned spins = 0;
    while (true) {
#pragma unroll
        for (int d = 0; d < PD; ++d) {
            bool ok = true;
#pragma unroll
            for (int i = 0; i < NWD; ++i) ok = ok && (i >= nw || (unsigned)(w[d][i] >> 32) == seq);
            if (__all(ok)) {
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    if constexpr (W == 1) {
                        v[q] = q < nvalid ? __uint_as_float((unsigned)w[d][q]) : (T)0;
                    } else {
                        // the layout of Tag<double>::put: low word, then high word
                        const u64 lo = w[d][q * W], hi = w[d][q * W + 1];
                        v[q] = q < nvalid ? __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffull)))
                                          : (T)0;
                    }
                }
                return true;
            }
#pragma unroll
            for (int i = 0; i < NWD; ++i) w[d][i] = ld_agent(p + (i < nw ? i : 0));
        }
        ++spins;
        if (spins > spin_limit) {
            // -1: a chained triangular solve timed out; info != 0 stops every
            // kernel queued after this one
            if ((threadIdx.x & 63) == 0) {
                __hip_atomic_store(abort, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (info) __hip_atomic_store(info, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            return false;
        }
        if ((spins & 255u) == 0 &&
            __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            return false;
        __builtin_amdgcn_s_sleep(1);
    }
}

template <typename T, int QB, bool LOWER, bool UNIT>
__global__ __launch_bounds__(NTC) void trsv_chain_kernel(const T *__restrict__ A, long lda, T *X, long ldx, int n,
                                                         int nrhs, u64 *xch, unsigned seq, int *ticket, int *abort,
                                                         unsigned spin_limit, int *info,
                                                         const T *__restrict__ dinv, const T *__restrict__ dmt)
{
    if (info && *info != 0) return;
    // blockIdx.y: a group of QB right-hand sides, an independent chain with its
    // own ticket and exchange region (the groups run side by side)
    const int grp = blockIdx.y;
    X += (size_t)grp * QB;
    const int nr = min(QB, nrhs - grp * QB);
    xch += (size_t)grp * n * QB * Tag<T>::W;
    ticket += grp;
    constexpr int W = Tag<T>::W;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *Ld = reinterpret_cast<T *>(smem_raw);   // [TB][LDP] diagonal block
    T *Lt = Ld + TB * LDP;                      // [TB][LDP] last off-diagonal tile
    T *part = Lt + TB * LDP;                    // [3][TB][QB] partial sums of waves 1-3
    __shared__ int s_ticket;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) s_ticket = atomicAdd(ticket, 1);
    __syncthreads();
    const int b = s_ticket;                      // position in dependency order
    const int nblk = (n + TB - 1) / TB;
    const int p = LOWER ? b : nblk - 1 - b;      // block row solved here
    const int r0 = p * TB, bs = min(TB, n - r0);
    bool failed = false;

    if (wave > 0) {
        // ---- off-diagonal tiles of dependencies 0 .. b-2, round robin over waves 1-3
        const int g = lane >> 4, c = (lane & 15) * 4;
        T acc[16][QB];
#pragma unroll
        for (int m = 0; m < 16; ++m)
#pragma unroll
            for (int q = 0; q < QB; ++q) acc[m][q] = (T)0;
        for (int j = wave - 1; j < b - 1; j += 3) {
            const int cb = LOWER ? j : nblk - 1 - j;
            const int c0 = cb * TB;
            T a[16][4];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int row = 4 * m + g;
                const T *src = A + (long)(r0 + row) * lda + c0 + c;
#pragma unroll
                for (int e = 0; e < 4; ++e) a[m][e] = (row < bs && c0 + c + e < n) ? src[e] : (T)0;
            }
            T xv[4][QB];
            bool ok = true;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int col = c0 + c + e;
                ok = ok && poll_vals<T, QB, 1>(xch + (size_t)col * QB * W, col < n ? QB : 0, seq, xv[e], spin_limit,
                                               abort, info);
            }
            if (!ok) { failed = true; break; }
#pragma unroll
            for (int m = 0; m < 16; ++m)
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int q = 0; q < QB; ++q) acc[m][q] += a[m][e] * xv[e][q];
        }
        // reduce over the 16 lanes that share a row; lane (m) of each group keeps row 4m+g
#pragma unroll
        for (int m = 0; m < 16; ++m)
#pragma unroll
            for (int q = 0; q < QB; ++q) {
                T v = acc[m][q];
#pragma unroll
                for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
                if ((lane & 15) == m) part[((wave - 1) * TB + 4 * m + g) * QB + q] = v;
            }
    } else {
        // ---- wave 0: stage the diagonal block (or its inverse) and the last
        //      dependency's tile (coalesced rows)
        if (dinv) {
            const T *Dp = dinv + (size_t)p * TB * TB;
            for (int i = 0; i < TB; ++i) Ld[i * LDP + lane] = Dp[i * TB + lane];
        } else {
            for (int i = 0; i < TB; ++i)
                Ld[i * LDP + lane] = (i < bs && lane < bs) ? A[(long)(r0 + i) * lda + r0 + lane] : (T)0;
        }
        if (b > 0) {
            const int cb = LOWER ? b - 1 : nblk - b;
            const int c0 = cb * TB, cs = min(TB, n - c0);
            if (dmt) {
                const T *Mp = dmt + (size_t)p * TB * TB;   // D^-1 times the last tile
                for (int i = 0; i < TB; ++i) Lt[i * LDP + lane] = Mp[i * TB + lane];
            } else {
                for (int i = 0; i < TB; ++i)
                    Lt[i * LDP + lane] = (i < bs && lane < cs) ? A[(long)(r0 + i) * lda + c0 + lane] : (T)0;
            }
        }
    }
    __syncthreads();
    if (wave != 0) return;

    // ---- critical path (wave 0, lane = row of the block)
    T s[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        s[q] = (lane < bs && q < nr) ? X[(long)(r0 + lane) * ldx + q] : (T)0;
        if (b > 1) s[q] -= part[lane * QB + q] + part[(TB + lane) * QB + q] + part[(2 * TB + lane) * QB + q];
    }
    if (dinv && dmt) {
        // x = D^-1 s - (D^-1 T) x_prev: the first product before the last
        // dependency's x arrives, only the second one after it
        T xo[QB];
#pragma unroll
        for (int q = 0; q < QB; ++q) xo[q] = (T)0;
#pragma unroll
        for (int k = 0; k < TB; ++k) {
            const T d = Ld[lane * LDP + k];
#pragma unroll
            for (int q = 0; q < QB; ++q) xo[q] += d * readlane(s[q], k);
        }
        if (b > 0) {
            const int cb = LOWER ? b - 1 : nblk - b;
            const int c0 = cb * TB;
            T xk[QB];
            if (!poll_vals<T, QB, 1>(xch + (size_t)(c0 + lane) * QB * W, c0 + lane < n ? QB : 0, seq, xk, spin_limit,
                                     abort, info))
                return;
#pragma unroll
            for (int k = 0; k < TB; ++k) {
                const T m = Lt[lane * LDP + k];
#pragma unroll
                for (int q = 0; q < QB; ++q) xo[q] -= m * readlane(xk[q], k);
            }
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) s[q] = xo[q];
    } else {
    if (b > 0) {
        const int cb = LOWER ? b - 1 : nblk - b;
        const int c0 = cb * TB;
        T xk[QB];
        if (!poll_vals<T, QB, 1>(xch + (size_t)(c0 + lane) * QB * W, c0 + lane < n ? QB : 0, seq, xk, spin_limit,
                                 abort, info))
            failed = true;
        if (failed) return;
#pragma unroll
        for (int k = 0; k < TB; ++k) {
            const T l = Lt[lane * LDP + k];
#pragma unroll
            for (int q = 0; q < QB; ++q) s[q] -= l * readlane(xk[q], k);
        }
    }
    if (failed) return;
    if (dinv) {
        // x = D^-1 s with the block's precomputed inverse: 64 independent
        // products instead of the 64-step dependent substitution
        T xo[QB];
#pragma unroll
        for (int q = 0; q < QB; ++q) xo[q] = (T)0;
#pragma unroll
        for (int k = 0; k < TB; ++k) {
            const T d = Ld[lane * LDP + k];
#pragma unroll
            for (int q = 0; q < QB; ++q) xo[q] += d * readlane(s[q], k);
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) s[q] = xo[q];
    } else {
    // FP64 divides by the diagonal as the reference does; FP32 multiplies by a
    // reciprocal each lane computes up front, which keeps the division off the
    // 64-step dependent chain (the FP32 factors only precondition refinement).
    constexpr bool EXACT = sizeof(T) == 8;
    const T rinv = (!UNIT && !EXACT && lane < bs) ? (T)1 / Ld[lane * LDP + lane] : (T)1;
    if (LOWER) {
#pragma unroll
        for (int k = 0; k < TB; ++k) {
            if (k < bs) {
                const T l = Ld[lane * LDP + k];
                const T d = Ld[k * LDP + k];
                const T rk = readlane(rinv, k);
#pragma unroll
                for (int q = 0; q < QB; ++q) {
                    T x = readlane(s[q], k);
                    if (!UNIT) x = EXACT ? x / d : x * rk;
                    if (lane == k) s[q] = x;
                    else if (lane > k) s[q] -= l * x;
                }
            }
        }
    } else {
#pragma unroll
        for (int k = TB - 1; k >= 0; --k) {
            if (k < bs) {
                const T l = Ld[lane * LDP + k];
                const T d = Ld[k * LDP + k];
                const T rk = readlane(rinv, k);
#pragma unroll
                for (int q = 0; q < QB; ++q) {
                    T x = readlane(s[q], k);
                    if (!UNIT) x = EXACT ? x / d : x * rk;
                    if (lane == k) s[q] = x;
                    else if (lane < k) s[q] -= l * x;
                }
            }
        }
    }
    }
    }
    if (lane < bs) {
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            if (q < nr) X[(long)(r0 + lane) * ldx + q] = s[q];
            Tag<T>::put(xch + ((size_t)(r0 + lane) * QB + q) * W, seq, s[q]);
        }
    }
}

// Two block rows per link: the FP32 single right-hand-side solve with the
// inverted diagonal blocks D_p^-1 and the couplings M_d[p] = D_p^-1 T(p, p-d)
// (d = 1, 2, 3 blocks back in dependency order).  Link q solves positions
// e = 2q and o = 2q + 1 (block rows p_e, p_o), which depend on the previous
// link's pair (e-1, e-2) and on their own streamed partial sums:
//   x_e = y_e - M_2[e] x_{e-2} - M_1[e] x_{e-1}
//   x_o = y_o - M_3[o] x_{e-2} - M_2[o] x_{e-1} - M_1[o] x_e
// with y = D^-1 (b - the products with the blocks before e-2).  Wave 0 finds
// x_e and publishes it at once (the next link's first product needs only it),
// wave 1 finds x_o, taking x_e from wave 0 through LDS; waves 2-7 stream the
// off-diagonal tiles, three per half.  One global hand-over per 128 rows
// instead of per 64: the hand-over (a store becoming visible to another CU,
// about 2 us) bounds the chain.
constexpr int NTP = 512;
constexpr int PLM = 7;   // LDS matrices: D_e^-1, M_2[e], M_1[e]; D_o^-1, M_3[o], M_2[o], M_1[o]

template <bool LOWER>
__global__ __launch_bounds__(NTP) void trsv_pair_kernel(const float *__restrict__ A, long lda, float *X, int n,
                                                        u64 *xch, unsigned seq, int *ticket, int *abort,
                                                        unsigned spin_limit, int *info,
                                                        const float *__restrict__ dinv, const float *__restrict__ dm)
{
    if (info && *info != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *Mx = reinterpret_cast<float *>(smem_raw);   // [PLM][TB][LDP]
    float *part = Mx + PLM * TB * LDP;                 // [6][TB] partial sums of waves 2-7
    // [TB] x_e for wave 1, each value with its own flag in the high word
    // (1: valid, 2: wave 0 gave up), so no fence orders value and flag
    u64 *xe_w = reinterpret_cast<u64 *>(part + 6 * TB);
    __shared__ int s_ticket;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) s_ticket = atomicAdd(ticket, 1);
    if (t < TB) xe_w[t] = 0ull;
    __syncthreads();
    const int q = s_ticket;
    const int nblk = (n + TB - 1) / TB;
    const int e = 2 * q, o = e + 1;
    const bool has_o = o < nblk;
    const size_t mstride = (size_t)nblk * TB * TB;   // between the couplings of distance d and d + 1
    auto blk = [&](int pos) { return LOWER ? pos : nblk - 1 - pos; };

    if (wave >= 2) {
        // ---- off-diagonal tiles of positions 0 .. e-3: half h = ws & 1, positions j = ws/2 mod 3
        // (a half past the last block row streams nothing, but still meets the barrier)
        const int ws = wave - 2, h = ws & 1;
        const int p = blk(e + h), r0 = p * TB, bs = min(TB, n - r0);
        const int jend = (h == 1 && !has_o) ? 0 : e - 2;
        const int g = lane >> 4, c = (lane & 15) * 4;
        float acc[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) acc[m] = 0.0f;
        bool failed = false;
        for (int j = ws >> 1; j < jend; j += 3) {
            const int c0 = blk(j) * TB;
            float a[16][4];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int row = 4 * m + g;
                const float *src = A + (long)(r0 + row) * lda + c0 + c;
#pragma unroll
                for (int u = 0; u < 4; ++u) a[m][u] = (row < bs && c0 + c + u < n) ? src[u] : 0.0f;
            }
            float xv[4];
            const int nv = min(max(n - (c0 + c), 0), 4);
            if (!poll_vals<float, 4, 2>(xch + (size_t)(c0 + c), nv, seq, xv, spin_limit, abort, info)) {
                failed = true;
                break;
            }
#pragma unroll
            for (int m = 0; m < 16; ++m)
#pragma unroll
                for (int u = 0; u < 4; ++u) acc[m] += a[m][u] * xv[u];
        }
        if (failed) {
#pragma unroll
            for (int m = 0; m < 16; ++m) acc[m] = 0.0f;   // the chain is abandoned; keep the barrier
        }
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            float v = acc[m];
#pragma unroll
            for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
            if ((lane & 15) == m) part[ws * TB + 4 * m + g] = v;
        }
    } else if (wave == 0 || has_o) {
        // ---- waves 0 and 1 stage their half's inverse and couplings (row i, column lane)
        const int h = wave, pos = e + h, p = blk(pos);
        float *D = Mx + (h == 0 ? 0 : 3) * TB * LDP;
        const float *Dp = dinv + (size_t)p * TB * TB;
        for (int i = 0; i < TB; ++i) D[i * LDP + lane] = Dp[i * TB + lane];
        const int nm = h == 0 ? 2 : 3;   // couplings to stage: d = 2, 1 (e) or d = 3, 2, 1 (o)
        for (int k = 0; k < nm; ++k) {
            const int d = nm - k;
            if (pos - d < 0) continue;   // no such dependency
            const float *Mp = dm + (size_t)(d - 1) * mstride + (size_t)p * TB * TB;
            float *Ml = D + (1 + k) * TB * LDP;
            for (int i = 0; i < TB; ++i) Ml[i * LDP + lane] = Mp[i * TB + lane];
        }
    }
    __syncthreads();
    if (wave >= 2 || (wave == 1 && !has_o)) return;

    // ---- critical path (wave h, lane = row of the block)
    const int h = wave, pos = e + h, p = blk(pos), r0 = p * TB, bs = min(TB, n - r0);
    const float *D = Mx + (h == 0 ? 0 : 3) * TB * LDP;
    float sv = lane < bs ? X[r0 + lane] : 0.0f;
    sv -= part[h * TB + lane] + part[(h + 2) * TB + lane] + part[(h + 4) * TB + lane];
    float y = 0.0f;
#pragma unroll
    for (int k = 0; k < TB; ++k) y += D[lane * LDP + k] * readlane(sv, k);
    // the coupling with the last dependency (x_{e-1} for wave 0, x_e for wave 1)
    // is applied after its x arrives: its row is held in registers beforehand
    float m1[TB];
    {
        const float *M1 = D + (h == 0 ? 2 : 3) * TB * LDP;
#pragma unroll
        for (int k = 0; k < TB; ++k) m1[k] = M1[lane * LDP + k];
#pragma unroll
        for (int k = 0; k < TB; ++k) asm volatile("" : "+v"(m1[k]));
    }
    bool ok = true;
    if (q > 0) {
        // the previous link's pair: x_{e-2} first (published first), then x_{e-1}
        float xk[1];
        int c0 = blk(e - 2) * TB;
        ok = poll_vals<float, 1, 4>(xch + (size_t)(c0 + lane), c0 + lane < n ? 1 : 0, seq, xk, spin_limit, abort, info);
        if (ok) {
            const float *Ml = D + TB * LDP;   // M_2[e] or M_3[o]
#pragma unroll
            for (int k = 0; k < TB; ++k) y -= Ml[lane * LDP + k] * readlane(xk[0], k);
            c0 = blk(e - 1) * TB;
            ok = poll_vals<float, 1, 4>(xch + (size_t)(c0 + lane), c0 + lane < n ? 1 : 0, seq, xk, spin_limit, abort,
                                        info);
        }
        if (ok) {
            if (h == 0) {
#pragma unroll
                for (int k = 0; k < TB; ++k) y -= m1[k] * readlane(xk[0], k);
            } else {
                const float *Ml = D + 2 * TB * LDP;   // M_2[o]
#pragma unroll
                for (int k = 0; k < TB; ++k) y -= Ml[lane * LDP + k] * readlane(xk[0], k);
            }
        }
    }
    if (h == 0) {
        if (ok && lane < bs) {
            X[r0 + lane] = y;
            Tag<float>::put(xch + r0 + lane, seq, y);
        }
        // every path hands wave 1 its words (flag 2: give up)
        __hip_atomic_store(&xe_w[lane], ((u64)(ok ? 1u : 2u) << 32) | __float_as_uint(y), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
        return;
    }
    if (!ok) return;
    // wave 1: x_e from wave 0 (a bounded wait: wave 0 writes its words on every path)
    u64 wv = 0;
    unsigned spins = 0;
    while (true) {
        wv = __hip_atomic_load(&xe_w[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (__all((wv >> 32) != 0ull)) break;
        if (++spins > 64u * spin_limit) {
            if (lane == 0) {
                __hip_atomic_store(abort, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (info) __hip_atomic_store(info, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    if (__any((wv >> 32) != 1ull)) return;
    const float xe = __uint_as_float((unsigned)wv);
#pragma unroll
    for (int k = 0; k < TB; ++k) y -= m1[k] * readlane(xe, k);
    if (lane < bs) {
        X[r0 + lane] = y;
        Tag<float>::put(xch + r0 + lane, seq, y);
    }
}

template <typename T, int QB, bool LOWER, bool UNIT>
int launch_chain(int n, int nr, const T *A, long lda, T *X, long ldx, ScirsWork &w, hipStream_t s, const T *dinv,
                 const T *dmt)
{
    const int groups = (nr + QB - 1) / QB;
    const auto kern = trsv_chain_kernel<T, QB, LOWER, UNIT>;
    const size_t lds = (2 * (size_t)TB * LDP + 3 * (size_t)TB * QB) * sizeof(T);
    if (lds > 64 * 1024) {
        int rc = scirs_allow_lds((const void *)kern, lds);
        if (rc) return rc;
    }
    int rc = scirs_fill_int(w.ticket, groups, 0, s);
    if (rc) return rc;
    unsigned seq = 0;
    if ((rc = scirs_next_seq(w, 512u, s, &seq))) return rc;
    const int nblk = (n + TB - 1) / TB;
    hipLaunchKernelGGL(kern, dim3(nblk, groups), dim3(NTC), lds, s, A, lda, X, ldx, n, nr, w.xtri, seq, w.ticket,
                       w.abort, w.spin_limit, w.info, dinv, dmt);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

template <typename T, int QB>
int chain_dispatch(int n, int nr, const T *A, long lda, T *X, long ldx, int lower, int unit, ScirsWork &w,
                   hipStream_t s, const T *dinv, const T *dmt)
{
    if (lower && unit) return launch_chain<T, QB, true, true>(n, nr, A, lda, X, ldx, w, s, dinv, dmt);
    if (lower) return launch_chain<T, QB, true, false>(n, nr, A, lda, X, ldx, w, s, dinv, dmt);
    if (unit) return launch_chain<T, QB, false, true>(n, nr, A, lda, X, ldx, w, s, dinv, dmt);
    return launch_chain<T, QB, false, false>(n, nr, A, lda, X, ldx, w, s, dinv, dmt);
}

int launch_pair(int n, const float *A, long lda, float *X, ScirsWork &w, hipStream_t s, bool lower,
                const float *dinv, const float *dm)
{
    const auto kern = lower ? trsv_pair_kernel<true> : trsv_pair_kernel<false>;
    const size_t lds = ((size_t)PLM * TB * LDP + 8 * (size_t)TB) * sizeof(float);
    int rc = scirs_allow_lds((const void *)kern, lds);
    if (rc) return rc;
    if ((rc = scirs_fill_int(w.ticket, 1, 0, s))) return rc;
    unsigned seq = 0;
    if ((rc = scirs_next_seq(w, 512u, s, &seq))) return rc;
    const int nblk = (n + TB - 1) / TB;
    hipLaunchKernelGGL(kern, dim3((nblk + 1) / 2), dim3(NTP), lds, s, A, lda, X, n, w.xtri, seq, w.ticket, w.abort,
                       w.spin_limit, w.info, dinv, dm);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

} // namespace

template <typename T>
int scirs_trsv_chain(int n, int nrhs, const T *A, long lda, T *X, long ldx, int lower, int unit, ScirsWork &w,
                     hipStream_t s, const T *dinv, const T *dmt)
{
    if (n <= 0 || nrhs <= 0) return SCIRS_OK;
    // right-hand sides per chain; up to SCIRS_CHAIN_GROUPS chains run side by side in one launch
    // (FP64 chains of four measured slower than chains of two: the n = 8192,
    // 16-rhs LU solve 52 -> 56 ms, the Cholesky solve 20 -> 25 ms)
    constexpr int QMAX = sizeof(T) == 8 ? 2 : 4;
    const int qb = nrhs == 1 ? 1 : (nrhs == 2 || QMAX == 2) ? 2 : 4;
    const int per_launch = qb * SCIRS_CHAIN_GROUPS;
    const size_t words = (size_t)n * std::min(nrhs, per_launch) * Tag<T>::W + (size_t)n * qb * Tag<T>::W;
    if (words > w.xtri_words)
        if (int rc = scirs::ctx().ensure_xtri((int)std::min<size_t>((words + 7) / 8, 1u << 30))) return rc;
    if (words > w.xtri_words) {
        scirs_set_error("chained triangular solve: exchange buffer holds %zu words, needs %zu", w.xtri_words, words);
        return SCIRS_ERR_ARG;
    }
    scirs::ProfScope prof(SCIRS_PROF_TRSV, (double)n * n * nrhs, (0.5 * n * n + 2.0 * n * nrhs) * sizeof(T), s);
    if constexpr (sizeof(T) == 4) {
        if (nrhs == 1 && dinv && dmt && w.trsv_pair && n > 2 * TB)
            return launch_pair(n, A, lda, X, w, s, lower != 0, dinv, dmt);
    }
    for (int rb = 0; rb < nrhs; rb += per_launch) {
        const int nr = std::min(per_launch, nrhs - rb);
        int rc;
        if (qb == 1) rc = chain_dispatch<T, 1>(n, nr, A, lda, X + rb, ldx, lower, unit, w, s, dinv, dmt);
        else if (qb == 2) rc = chain_dispatch<T, 2>(n, nr, A, lda, X + rb, ldx, lower, unit, w, s, dinv, dmt);
        else rc = chain_dispatch<T, QMAX>(n, nr, A, lda, X + rb, ldx, lower, unit, w, s, dinv, dmt);
        if (rc) return rc;
    }
    return SCIRS_OK;
}

// Inverses of the 64 x 64 diagonal blocks of a triangle (FP32 factors that
// only precondition refinement): Dinv[p] (row-major 64 x 64) with
// Dinv[p] * T_pp = I.  One wave per block, lane = column c of the inverse,
// found by substitution on e_c in registers (the block rows broadcast from
// LDS); rows past n are the identity.
namespace {
template <bool LOWER, bool UNIT>
__global__ __launch_bounds__(64) void tri_block_inverse_kernel(int n, const float *__restrict__ A, long lda,
                                                               float *__restrict__ dinv)
{
    __shared__ float D[TB][TB + 1];
    const int p = blockIdx.x, lane = threadIdx.x;
    const int r0 = p * TB, bs = min(TB, n - r0);
    for (int i = 0; i < TB; ++i)
        D[i][lane] = (i < bs && lane < bs) ? A[(long)(r0 + i) * lda + r0 + lane] : (i == lane ? 1.0f : 0.0f);
    __syncthreads();
    float y[TB];
    if (LOWER) {
#pragma unroll
        for (int i = 0; i < TB; ++i) {
            float v = (i == lane) ? 1.0f : 0.0f;
#pragma unroll
            for (int j = 0; j < i; ++j) v -= D[i][j] * y[j];
            y[i] = UNIT ? v : v / D[i][i];
        }
    } else {
#pragma unroll
        for (int i = TB - 1; i >= 0; --i) {
            float v = (i == lane) ? 1.0f : 0.0f;
#pragma unroll
            for (int j = i + 1; j < TB; ++j) v -= D[i][j] * y[j];
            y[i] = UNIT ? v : v / D[i][i];
        }
    }
    float *Dp = dinv + (size_t)p * TB * TB;
#pragma unroll
    for (int i = 0; i < TB; ++i) Dp[i * TB + lane] = y[i];
}
} // namespace

int scirs_tri_block_inverse(int n, const float *A, long lda, int lower, int unit, float *dinv, hipStream_t s)
{
    if (n <= 0) return SCIRS_OK;
    const int nblk = (n + TB - 1) / TB;
    if (lower && unit) hipLaunchKernelGGL((tri_block_inverse_kernel<true, true>), dim3(nblk), dim3(64), 0, s, n, A, lda, dinv);
    else if (lower) hipLaunchKernelGGL((tri_block_inverse_kernel<true, false>), dim3(nblk), dim3(64), 0, s, n, A, lda, dinv);
    else if (unit) hipLaunchKernelGGL((tri_block_inverse_kernel<false, true>), dim3(nblk), dim3(64), 0, s, n, A, lda, dinv);
    else hipLaunchKernelGGL((tri_block_inverse_kernel<false, false>), dim3(nblk), dim3(64), 0, s, n, A, lda, dinv);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

// M_d[p] = Dinv[p] * T(p, p -+ d) for the tile block row p meets d-th from
// last in the chained solve (block column p-d of a lower triangle, p+d of an
// upper one), d = 1 .. depth, so a link applies D^-1 to its partial sums
// before those dependencies' x arrive and only M_d[p] x after.  One workgroup
// per (block row, d); M_d at dmt + (d-1) * ceil(n/64) * 64 * 64.
namespace {
template <bool LOWER>
__global__ __launch_bounds__(256) void tri_block_coupling_kernel(int n, const float *__restrict__ A, long lda,
                                                                 const float *__restrict__ dinv,
                                                                 float *__restrict__ dmt)
{
    __shared__ float Dl[TB][TB + 1], Tl[TB][TB + 1];
    const int p = blockIdx.x, nblk = (n + TB - 1) / TB, d = blockIdx.y + 1;
    const int cb = LOWER ? p - d : p + d;
    if (cb < 0 || cb >= nblk) return;
    dmt += (size_t)blockIdx.y * nblk * TB * TB;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r0 = p * TB, bs = min(TB, n - r0), c0 = cb * TB, cs = min(TB, n - c0);
    const float *Dp = dinv + (size_t)p * TB * TB;
    for (int i = wave; i < TB; i += 4) {
        Dl[i][lane] = Dp[i * TB + lane];
        Tl[i][lane] = (i < bs && lane < cs) ? A[(long)(r0 + i) * lda + c0 + lane] : 0.0f;
    }
    __syncthreads();
    float *Mp = dmt + (size_t)p * TB * TB;
    for (int i = wave; i < TB; i += 4) {
        float acc = 0.0f;
#pragma unroll 16
        for (int k = 0; k < TB; ++k) acc += Dl[i][k] * Tl[k][lane];
        Mp[i * TB + lane] = acc;
    }
}
} // namespace

int scirs_tri_block_coupling(int n, const float *A, long lda, int lower, const float *dinv, float *dmt, int depth,
                             hipStream_t s)
{
    if (n <= 0) return SCIRS_OK;
    if (depth < 1 || depth > 3) {
        scirs_set_error("tri_block_coupling: depth %d outside 1..3", depth);
        return SCIRS_ERR_ARG;
    }
    const int nblk = (n + TB - 1) / TB;
    const dim3 grid(nblk, depth);
    if (lower) hipLaunchKernelGGL(tri_block_coupling_kernel<true>, grid, dim3(256), 0, s, n, A, lda, dinv, dmt);
    else hipLaunchKernelGGL(tri_block_coupling_kernel<false>, grid, dim3(256), 0, s, n, A, lda, dinv, dmt);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

// U = L^T for the lower triangle of L (64 x 64 tiles through LDS; only the tiles
// on or above U's diagonal are written).  Turns the Cholesky factor into the
// row-major upper triangle the chained solve streams by rows.
namespace {
template <typename T>
__global__ __launch_bounds__(NTC) void transpose_lower_kernel(int n, const T *__restrict__ L, long ldl, T *U, long ldu)
{
    __shared__ T tile[TB][TB + 1];
    // blockIdx.x enumerates the lower-triangle tiles (bi >= bj) row by row
    const int idx = blockIdx.x;
    int bi = (int)((sqrtf(8.0f * (float)idx + 1.0f) - 1.0f) * 0.5f);
    while (bi * (bi + 1) / 2 > idx) --bi;
    while ((bi + 1) * (bi + 2) / 2 <= idx) ++bi;
    const int bj = idx - bi * (bi + 1) / 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = bi * TB, c0 = bj * TB;
    for (int i = wave; i < TB; i += NTC / 64) {
        const int r = r0 + i, c = c0 + lane;
        tile[i][lane] = (r < n && c < n) ? L[(long)r * ldl + c] : (T)0;
    }
    __syncthreads();
    // U[c0 + i][r0 + lane] = L[r0 + lane][c0 + i]
    for (int i = wave; i < TB; i += NTC / 64) {
        const int r = c0 + i, c = r0 + lane;
        if (r < n && c < n) U[(long)r * ldu + c] = tile[lane][i];
    }
}
} // namespace

template <typename T>
int scirs_transpose_lower(int n, const T *L, long ldl, T *U, long ldu, hipStream_t s)
{
    if (n <= 0) return SCIRS_OK;
    const long nb = (n + TB - 1) / TB;
    const long tiles = nb * (nb + 1) / 2;
    hipLaunchKernelGGL(transpose_lower_kernel<T>, dim3((unsigned)tiles), dim3(NTC), 0, s, n, L, ldl, U, ldu);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}
template int scirs_transpose_lower<double>(int, const double *, long, double *, long, hipStream_t);
template int scirs_transpose_lower<float>(int, const float *, long, float *, long, hipStream_t);

template int scirs_trsv_chain<float>(int, int, const float *, long, float *, long, int, int, ScirsWork &, hipStream_t,
                                     const float *, const float *);
template int scirs_trsv_chain<double>(int, int, const double *, long, double *, long, int, int, ScirsWork &,
                                      hipStream_t, const double *, const double *);
