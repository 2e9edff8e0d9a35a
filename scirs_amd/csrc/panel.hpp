// Cooperative partial-pivot panel factorisation (included by getrf.hip).
//
// The panel's rows are split into contiguous chunks, one per workgroup, and
// stay in LDS for the whole panel.  Per column j every workgroup publishes
// its candidate (largest |a_ij| of its chunk, lowest row on ties) and that
// row's full contents; every workgroup then reads all candidates and picks the
// same winner — the strict '>' scan of scirs2-linalg/src/lapack.rs:95-104 —
// swaps, divides (lapack.rs:138) and updates its own rows (lapack.rs:140-142).
//
// Exchange protocol: every 64-bit word written to the exchange buffers
// carries the 32-bit sequence number of its column in the high half (a value
// wider than 32 bits travels as several tagged words), so a reader knows a
// word is current without any fence or atomic read-modify-write: relaxed
// agent-scope atomic stores and loads only.  Readers poll until every word
// they need carries the expected number; polls are bounded and raise *abort
// when exceeded, so no wave can wait forever.
//
// Latency hiding: (1) every poll of the candidate list also reads the row of
// the best candidate of the previous poll (in the same round trip, and kept
// once complete), so when the last candidate arrives the pivot row is usually
// already in registers — one memory round trip per column instead of two;
// (2) right after the pivot row arrives, a workgroup computes the
// multipliers, updates column j+1 and finds its next candidate in one pass,
// brings that row (and row j+1 when it holds it) fully up to date and
// publishes; (3) the bulk rank-1 update of its other rows runs on waves 1-3
// while wave 0 already polls the next column.
//
// A panel whose rows fit one workgroup's LDS (small matrices, the last panels)
// runs as a single 512-thread workgroup: its local arg-max is the global one,
// so nothing is published or polled, and the pivot row is swapped in place.
#pragma once
#include "common.hpp"

namespace panel {

constexpr int NT = 256;
constexpr int SCIRS_PANEL_KB_MAX = 256;   // widest panel (the launcher rejects wider)
typedef unsigned long long u64;

__device__ __forceinline__ void st_agent(u64 *p, u64 v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 ld_agent(const u64 *p)
{
    return __hip_atomic_load(const_cast<u64 *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T> struct Tag;
template <> struct Tag<float> {
    static constexpr int W = 1;   // words per value
    __device__ static void put(u64 *dst, unsigned seq, float v)
    {
        st_agent(dst, ((u64)seq << 32) | __float_as_uint(v));
    }
    __device__ static bool get(const u64 *src, unsigned seq, float &v)
    {
        const u64 w = ld_agent(src);
        v = __uint_as_float((unsigned)w);
        return (unsigned)(w >> 32) == seq;
    }
    __device__ static void put_s(u64 *dst, size_t, unsigned seq, float v) { put(dst, seq, v); }
    __device__ static bool get_s(const u64 *src, size_t, unsigned seq, float &v) { return get(src, seq, v); }
};
template <> struct Tag<double> {
    static constexpr int W = 2;
    __device__ static void put(u64 *dst, unsigned seq, double v) { put_s(dst, 1, seq, v); }
    __device__ static bool get(const u64 *src, unsigned seq, double &v) { return get_s(src, 1, seq, v); }
    // the two words `stride` apart
    __device__ static void put_s(u64 *dst, size_t stride, unsigned seq, double v)
    {
        const u64 bits = (u64)__double_as_longlong(v);
        st_agent(dst, ((u64)seq << 32) | (bits & 0xffffffffull));
        st_agent(dst + stride, ((u64)seq << 32) | (bits >> 32));
    }
    __device__ static bool get_s(const u64 *src, size_t stride, unsigned seq, double &v)
    {
        const u64 lo = ld_agent(src), hi = ld_agent(src + stride);
        v = __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffull)));
        return (unsigned)(lo >> 32) == seq && (unsigned)(hi >> 32) == seq;
    }
};

template <typename T>
struct PanelArgs {
    T *A; long lda;
    int k0, m, kb, R, G;
    u64 *xw;       // [nrep][2][4][G]   candidate: word 0 (seq|row), then the value words;
                   //                   word k of workgroup q at [k][q], so one wave's poll of
                   //                   all candidates reads consecutive words
    u64 *xrow;     // [nrep][2][G][kb*W] candidate rows
    u64 *xrj;      // [nrep][2][kb*W]   row j of the panel
    int nrep;      // replicas of every published word; workgroup g reads replica g % nrep
    unsigned seq_base;
    int *pairs; int *npairs;
    int *info; double *info_val; int *abort;
    unsigned spin_limit;
    int poll_sleep;   // s_sleep(2) repetitions after an unsuccessful candidate poll
};

// Rank-1 update row -= lcol[row] * prow over rows [r1, R) except skip0/skip1
// and columns [qlo, kb), by nthr threads (tt = 0..nthr-1, nthr a multiple of
// 32).  32 consecutive threads cover 32 consecutive columns of one row
// (conflict-free LDS rows), each thread KMAX strided columns whose prow values
// stay in registers; RU rows are loaded before any is written back, so every
// thread keeps RU * KMAX independent LDS reads in flight (one wave per SIMD
// cannot hide LDS latency any other way).  The code is branch-free: an element
// outside the update (a skipped row, a row past R, a column past kb) is read
// from and written back to its row's padding column kb, which holds no data.
template <typename T, int KMAX>
__device__ __forceinline__ void bulk_rank1(T *tile, int kbp, int kb, const T *lcol, const T *prow, int r1, int R,
                                           int qlo, int skip0, int skip1, int tt, int nthr)
{
    constexpr int TPR = 32, RU = 4;
    if (qlo >= kb || r1 >= R) return;
    const int qt = tt % TPR, rstep = nthr / TPR;
    T pr[KMAX];
    int qc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const int q = qlo + qt + TPR * k;
        qc[k] = q < kb ? q : kb;
        pr[k] = prow[min(q, kb - 1)];
    }
    for (int r = r1 + tt / TPR; r < R; r += RU * rstep) {
        T v[RU][KMAX], l[RU];
        int ad[RU][KMAX];
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const int rr = r + u * rstep;
            const bool ok = rr < R && rr != skip0 && rr != skip1;
            const int rc = min(rr, R - 1);
            l[u] = lcol[rc];
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                ad[u][k] = rc * kbp + (ok ? qc[k] : kb);
                v[u][k] = tile[ad[u][k]];
            }
        }
#pragma unroll
        for (int u = 0; u < RU; ++u)
#pragma unroll
            for (int k = 0; k < KMAX; ++k) tile[ad[u][k]] = v[u][k] - l[u] * pr[k];
    }
}

template <typename T>
__device__ __forceinline__ void bulk_update(T *tile, int kbp, int kb, const T *lcol, const T *prow, int r1, int R,
                                            int qlo, int skip0, int skip1, int tt, int nthr)
{
    if (kb - qlo <= 64)
        bulk_rank1<T, 2>(tile, kbp, kb, lcol, prow, r1, R, qlo, skip0, skip1, tt, nthr);
    else if (kb - qlo <= 128)
        bulk_rank1<T, 4>(tile, kbp, kb, lcol, prow, r1, R, qlo, skip0, skip1, tt, nthr);
    else
        bulk_rank1<T, SCIRS_PANEL_KB_MAX / 32>(tile, kbp, kb, lcol, prow, r1, R, qlo, skip0, skip1, tt, nthr);
}

// W0 selects where the column's critical path runs: false = all four waves
// (five barriers per column), true = wave 0 alone (rows strided over its 64
// lanes; two barriers per column), waves 1-3 doing only the bulk update.
// NTK: threads per workgroup (256; a single-workgroup panel, with no
// exchange to overlap, uses 512)
template <typename T, bool W0, int NTK = NT>
__global__ __launch_bounds__(NTK) void getf2_panel_kernel(PanelArgs<T> a)
{
    if (*a.info != 0) return;
    constexpr int W = Tag<T>::W;
    constexpr int SPEC = SCIRS_PANEL_KB_MAX / 64;   // pivot-row values per lane of wave 0
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int kb = a.kb, kbp = kb + 1;
    T *tile = reinterpret_cast<T *>(smem_raw);        // R x kbp
    T *prow2 = tile + (size_t)a.R * kbp;              // [2][kb] pivot row, by column parity
    T *rjv = prow2 + 2 * kb;                          // kb  row j (for the owner of the pivot row)
    T *lcol = rjv + kb;                               // R   multipliers
    int *top_label = reinterpret_cast<int *>(lcol + a.R);   // kb (workgroup 0)
    int *out_pos = top_label + kb;                    // kb
    int *out_label = out_pos + kb;                    // kb
    __shared__ T red_v[NTK / 64];
    __shared__ int red_i[NTK / 64];
    __shared__ int s_p, s_g, s_stop, s_c, s_nout;
    __shared__ T s_pv, s_cv;

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int g = blockIdx.x, G = a.G;
    const int row0 = g * a.R;
    const int Rg = max(0, min(a.R, a.m - row0));
    T *Ap = a.A + (long)a.k0 * a.lda + a.k0;
    // Every published word is written to nrep replicas and workgroup g reads
    // replica g % nrep (with round-robin placement, its XCD's copy): the polls
    // of G workgroups spread over nrep times more cache lines.
    const int nrep = a.nrep, myrep = g % nrep;
    const size_t rs_w = (size_t)2 * G * 4, rs_r = (size_t)2 * G * kb * W, rs_j = (size_t)2 * kb * W;
    auto XW = [&](int par, int q, int r) { return a.xw + r * rs_w + (size_t)par * 4 * G + q; };   // word k at + k*G
    auto XR = [&](int par, int q, int r) { return a.xrow + r * rs_r + ((size_t)par * G + q) * (size_t)kb * W; };
    auto XJ = [&](int par, int r) { return a.xrj + r * rs_j + (size_t)par * kb * W; };

    g2s_block<T, 16>(tile, kbp, Ap + (long)row0 * a.lda, a.lda, Rg, kb, Rg, kb);   // 16 loads in flight per thread
    if (g == 0) {
        for (int q = t; q < kb; q += NTK) top_label[q] = q;
        if (t == 0) s_nout = 0;
    }
    if (t == 0) s_stop = 0;
    __syncthreads();

    // local arg-max of column `col` over local rows >= rstart -> s_c (local row, -1 none), s_cv
    auto local_candidate = [&](int col, int rstart) {
        T bv = (T)0; int bi = -1;
        for (int r = rstart + t; r < Rg; r += NTK) {
            const T v = fabs(tile[r * kbp + col]);
            if (bi < 0 || v > bv) { bv = v; bi = r; }
        }
        wave_argmax(bv, bi);
        if (lane == 0) { red_v[wave] = bv; red_i[wave] = bi; }
        __syncthreads();
        if (t == 0) {
            T v = red_v[0]; int i = red_i[0];
            for (int w = 1; w < NTK / 64; ++w) {
                const int oi = red_i[w]; const T ov = red_v[w];
                if (oi >= 0 && (i < 0 || ov > v || (ov == v && oi < i))) { v = ov; i = oi; }
            }
            s_c = i; s_cv = v;
        }
        __syncthreads();
    };
    // publish the candidate of column j (its row is fully up to date for column j)
    auto publish = [&](int j) {
        const int b = j & 1;
        const unsigned seq = a.seq_base + (unsigned)j + 1u;
        const int c = s_c;
        for (int r = 0; r < nrep; ++r) {
            u64 *xr = XR(b, g, r);
            if (c >= 0)
                for (int q = t; q < kb; q += NTK) Tag<T>::put(xr + (size_t)q * W, seq, tile[c * kbp + q]);
            if (j >= row0 && j < row0 + Rg) {
                u64 *xj = XJ(b, r);
                for (int q = t; q < kb; q += NTK) Tag<T>::put(xj + (size_t)q * W, seq, tile[(j - row0) * kbp + q]);
            }
            if (t == 0) {
                u64 *xw = XW(b, g, r);
                st_agent(xw, ((u64)seq << 32) | (unsigned)(c >= 0 ? row0 + c : -1));
                Tag<T>::put_s(xw + G, G, seq, s_cv);
            }
        }
    };

    local_candidate(0, 0);
    // One workgroup holds the whole panel: its local arg-max is the global one,
    // and the candidate never leaves the workgroup (no publish, no poll).
    const bool solo = !W0 && G == 1;
    int nc = s_c;       // solo: candidate row of the next column (local = global)
    T ncv = s_cv;
    if (!solo) publish(0);

    const T eps = num_limits<T>::eps();
    for (int j = 0; j < kb; ++j) {
        const int b = j & 1;
        const unsigned seq = a.seq_base + (unsigned)j + 1u;
        // solo: the pivot row is swapped into tile row j, which no later column changes
        T *prow = solo ? tile + (size_t)j * kbp : prow2 + b * kb;
        // ---- 1. wave 0 reduces the candidate list (the same in every workgroup) and,
        //         in the same polls, fetches the current best candidate's row and row j,
        //         so the pivot row usually arrives with the last candidate.  Waves 1-3
        //         are meanwhile finishing the previous column's bulk update.  (More
        //         polling waves were measured slower: they add contention on the
        //         candidate lines.)
        if (solo) {
            // every thread holds the candidate (nc, ncv): nothing to exchange
            if (t == 0 && !(nc >= 0 && ncv >= eps)) {   // |pivot| < eps: singular (lapack.rs:106-120)
                s_stop = 1;
                *a.info = a.k0 + j + 1;
                *a.info_val = (double)(nc >= 0 ? ncv : (T)0);
            }
        } else if (wave == 0) {
            bool mine = false;
            unsigned spins = 0;
            T vv = (T)0; int ii = -1, wgw = -1;
            bool timed_out = false, row_ok = false;
            T sp[SPEC], sj[SPEC];
            const u64 *xj = XJ(b, myrep);
            int guess = -1;   // best workgroup of the previous poll: its row is read with this poll's records
            bool have = false;   // sp (and sj) already hold the guess's complete row: not read again
            while (true) {
                bool ok = true;
                T v = (T)0; int i = -1, wg = -1;
                const bool fetch = guess >= 0 && !have;
                bool rok = fetch || have;
                const u64 *xr = XR(b, guess >= 0 ? guess : 0, myrep);
                // row j is needed only by the workgroup that holds the pivot row
                const bool want_j = guess == g;
                if (fetch) {
#pragma unroll
                    for (int u = 0; u < SPEC; ++u) {
                        const int q = lane + 64 * u;
                        if (q < kb) {
                            rok = Tag<T>::get(xr + (size_t)q * W, seq, sp[u]) && rok;
                            if (want_j) rok = Tag<T>::get(xj + (size_t)q * W, seq, sj[u]) && rok;
                        }
                    }
                }
                for (int q = lane; q < G; q += 64) {
                    const u64 *xw = XW(b, q, myrep);
                    const u64 w0 = ld_agent(xw);
                    T qv;
                    const bool okv = Tag<T>::get_s(xw + G, G, seq, qv);
                    ok = ok && okv && (unsigned)(w0 >> 32) == seq;
                    const int qi = (int)(unsigned)w0;
                    if (qi >= 0 && (i < 0 || qv > v || (qv == v && qi < i))) { v = qv; i = qi; wg = q; }
                }
                vv = v; ii = i;
                wave_argmax(vv, ii);
                const unsigned long long ball = __ballot(i == ii && i >= 0);
                const int src = ball ? __ffsll(ball) - 1 : 0;
                wgw = __builtin_amdgcn_readlane(wg, src);
                if (__all(ok)) { row_ok = wgw >= 0 && wgw == guess && __all(rok); mine = true; break; }
                if (wgw != guess) { guess = wgw; have = false; }
                else if (fetch && __all(rok)) have = true;
                for (int z = 0; z < a.poll_sleep; ++z) __builtin_amdgcn_s_sleep(2);
                if (++spins > a.spin_limit ||
                    __hip_atomic_load(a.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    timed_out = true;
                    mine = true;
                    break;
                }
            }
            // the guess missed: read the winner's row (and row j) now
            if (mine && !row_ok && !timed_out && wgw >= 0 && vv >= eps) {
                const u64 *xr = XR(b, wgw, myrep);
                const bool want_j = wgw == g;
                while (true) {
                    bool rok = true;
#pragma unroll
                    for (int u = 0; u < SPEC; ++u) {
                        const int q = lane + 64 * u;
                        if (q < kb) {
                            rok = Tag<T>::get(xr + (size_t)q * W, seq, sp[u]) && rok;
                            if (want_j) rok = Tag<T>::get(xj + (size_t)q * W, seq, sj[u]) && rok;
                        }
                    }
                    if (__all(rok)) { row_ok = true; break; }
                    if (++spins > a.spin_limit) { timed_out = true; break; }
                }
            }
            if (mine && row_ok) {
                const bool have_j = wgw == g;
#pragma unroll
                for (int u = 0; u < SPEC; ++u) {
                    const int q = lane + 64 * u;
                    if (q < kb) {
                        prow[q] = sp[u];
                        if (have_j) rjv[q] = sj[u];
                    }
                }
            }
            if (mine && lane == 0) {
                s_p = ii; s_g = wgw; s_pv = vv;
                if (timed_out) {
                    // the 1-based column whose exchange timed out; info != 0 stops
                    // every kernel queued after this one
                    __hip_atomic_store(a.abort, a.k0 + j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(a.info, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_stop = 1;
                } else if (!(vv >= eps)) {   // |pivot| < eps: singular (lapack.rs:106-120)
                    s_stop = 1;
                    if (g == 0) { *a.info = a.k0 + j + 1; *a.info_val = (double)vv; }
                }
            }
        }
        __syncthreads();
        if (s_stop) break;
        if constexpr (W0) {
        const int p = s_p;
        const bool own_j = (j >= row0 && j < row0 + Rg);
        const bool own_p = (p >= row0 && p < row0 + Rg);
        const int r1 = max(0, j + 1 - row0);
        const bool more = j + 1 < kb;
        if (wave == 0) {
            // ---- 2. swap, permutation record, multipliers, column j+1, arg-max: one wave
            if (p != j) {
                if (own_j)
                    for (int c = lane; c < kb; c += 64) tile[(j - row0) * kbp + c] = prow[c];
                if (own_p)
                    for (int c = lane; c < kb; c += 64) tile[(p - row0) * kbp + c] = rjv[c];
                if (g == 0) {
                    if (p < kb) {
                        if (lane == 0) { int tmp = top_label[j]; top_label[j] = top_label[p]; top_label[p] = tmp; }
                    } else {
                        const int nout = s_nout;
                        int found = -1;
                        for (int base = 0; base < nout && found < 0; base += 64) {
                            const int sidx = base + lane;
                            const unsigned long long bl = __ballot(sidx < nout && out_pos[sidx] == p);
                            if (bl) found = base + __ffsll(bl) - 1;
                        }
                        if (lane == 0) {
                            if (found < 0) { found = nout; out_pos[found] = p; out_label[found] = p; s_nout = nout + 1; }
                            int tmp = top_label[j]; top_label[j] = out_label[found]; out_label[found] = tmp;
                        }
                    }
                }
            }
            const T pivot = prow[j];   // true division (lapack.rs:138)
            const T u1 = more ? prow[j + 1] : (T)0;
            T bv = (T)0; int bi = -1;
            for (int r = r1 + lane; r < Rg; r += 64) {
                const T l = tile[r * kbp + j] / pivot;
                tile[r * kbp + j] = l;
                lcol[r] = l;
                if (more) {
                    const T x = tile[r * kbp + j + 1] - l * u1;
                    tile[r * kbp + j + 1] = x;
                    if (bi < 0 || fabs(x) > bv) { bv = fabs(x); bi = r; }
                }
            }
            wave_argmax(bv, bi);
            if (lane == 0) { s_c = bi; s_cv = bv; }
        }
        __syncthreads();
        if (!more) break;
        const int c = s_c;
        const int jn = (j + 1 >= row0 && j + 1 < row0 + Rg) ? j + 1 - row0 : -1;
        if (wave == 0) {
            // ---- 3. the next candidate row and row j+1 fully up to date, published
            for (int q = j + 2 + lane; q < kb; q += 64) {
                if (c >= 0) tile[c * kbp + q] -= lcol[c] * prow[q];
                if (jn >= 0 && jn != c) tile[jn * kbp + q] -= lcol[jn] * prow[q];
            }
            const int jj = j + 1, bb = jj & 1;
            const unsigned sq = a.seq_base + (unsigned)jj + 1u;
            for (int r = 0; r < nrep; ++r) {
                u64 *xr = XR(bb, g, r);
                if (c >= 0)
                    for (int q = lane; q < kb; q += 64) Tag<T>::put(xr + (size_t)q * W, sq, tile[c * kbp + q]);
                if (jn >= 0) {
                    u64 *xj = XJ(bb, r);
                    for (int q = lane; q < kb; q += 64) Tag<T>::put(xj + (size_t)q * W, sq, tile[jn * kbp + q]);
                }
                if (lane == 0) {
                    u64 *xw = XW(bb, g, r);
                    st_agent(xw, ((u64)sq << 32) | (unsigned)(c >= 0 ? row0 + c : -1));
                    Tag<T>::put_s(xw + G, G, sq, s_cv);
                }
            }
        } else if (j + 2 < kb) {
            // ---- 4. bulk rank-1 update of the other rows (waves 1-3), overlapping
            //         wave 0's publication and its next poll
            bulk_update<T>(tile, kbp, kb, lcol, prow, r1, Rg, j + 2, c, jn, t - 64, NTK - 64);
        }
        } else {
        const int p = solo ? nc : s_p;
        const bool own_j = (j >= row0 && j < row0 + Rg);
        const bool own_p = (p >= row0 && p < row0 + Rg);
        // ---- 3. row interchange inside the panel (the permutation record is kept
        //         later, off the critical path)
        if (solo && p != j) {
            for (int c = t; c < kb; c += NTK) {
                const T v = tile[j * kbp + c];
                tile[j * kbp + c] = tile[p * kbp + c];
                tile[p * kbp + c] = v;
            }
        } else if (p != j) {
            if (own_j)
                for (int c = t; c < kb; c += NTK) tile[(j - row0) * kbp + c] = prow[c];
            if (own_p)
                for (int c = t; c < kb; c += NTK) tile[(p - row0) * kbp + c] = rjv[c];
        }
        // workgroup 0 records row interchanges as labels (wave 0)
        auto record_swap = [&]() {
            if (p < kb) {
                if (lane == 0) { int tmp = top_label[j]; top_label[j] = top_label[p]; top_label[p] = tmp; }
            } else {
                const int nout = s_nout;
                int found = -1;
                for (int base = 0; base < nout && found < 0; base += 64) {
                    const int sidx = base + lane;
                    const unsigned long long bl = __ballot(sidx < nout && out_pos[sidx] == p);
                    if (bl) found = base + __ffsll(bl) - 1;
                }
                if (lane == 0) {
                    if (found < 0) { found = nout; out_pos[found] = p; out_label[found] = p; s_nout = nout + 1; }
                    int tmp = top_label[j]; top_label[j] = out_label[found]; out_label[found] = tmp;
                }
            }
        };
        __syncthreads();
        // ---- 4. multipliers by true division (lapack.rs:138), column j+1, and this
        //         thread's arg-max of column j+1, in one pass over the rows
        const T pivot = prow[j];
        const int r1 = max(0, j + 1 - row0);
        const bool more = j + 1 < kb;
        const T u1 = more ? prow[j + 1] : (T)0;
        T bv = (T)0; int bi = -1;
        for (int r = r1 + t; r < Rg; r += NTK) {
            const T l = tile[r * kbp + j] / pivot;
            tile[r * kbp + j] = l;
            lcol[r] = l;
            if (more) {
                const T x = tile[r * kbp + j + 1] - l * u1;
                tile[r * kbp + j + 1] = x;
                if (bi < 0 || fabs(x) > bv) { bv = fabs(x); bi = r; }
            }
        }
        if (!more) {
            if (g == 0 && wave == 0 && p != j) record_swap();
            break;
        }
        wave_argmax(bv, bi);
        if (lane == 0) { red_v[wave] = bv; red_i[wave] = bi; }
        __syncthreads();
        // every thread combines the four wave results itself (lowest row wins ties)
        T cv = red_v[0]; int c = red_i[0];
#pragma unroll
        for (int w = 1; w < NTK / 64; ++w) {
            const int oi = red_i[w]; const T ov = red_v[w];
            if (oi >= 0 && (c < 0 || ov > cv || (ov == cv && oi < c))) { cv = ov; c = oi; }
        }
        // ---- 5. the next candidate row and row j+1: each element brought up to
        //         date and published by the same thread (readers check every
        //         word's tag, so no barrier is needed before publishing)
        const int jn = (j + 1 >= row0 && j + 1 < row0 + Rg) ? j + 1 - row0 : -1;
        if (solo) {
            for (int q = j + 2 + t; q < kb; q += NTK) {
                if (c >= 0) tile[c * kbp + q] -= lcol[c] * prow[q];
                if (jn >= 0 && jn != c) tile[jn * kbp + q] -= lcol[jn] * prow[q];
            }
            nc = c;
            ncv = cv;
        } else {
            const int bb = (j + 1) & 1;
            const unsigned sq = a.seq_base + (unsigned)j + 2u;
            for (int q = t; q < kb; q += NTK) {
                if (c >= 0) {
                    T v = tile[c * kbp + q];
                    if (q >= j + 2) { v -= lcol[c] * prow[q]; tile[c * kbp + q] = v; }
                    for (int r = 0; r < nrep; ++r) Tag<T>::put(XR(bb, g, r) + (size_t)q * W, sq, v);
                }
                if (jn >= 0) {
                    T v = tile[jn * kbp + q];
                    if (q >= j + 2 && jn != c) { v -= lcol[jn] * prow[q]; tile[jn * kbp + q] = v; }
                    for (int r = 0; r < nrep; ++r) Tag<T>::put(XJ(bb, r) + (size_t)q * W, sq, v);
                }
            }
            if (t == 0) {
                for (int r = 0; r < nrep; ++r) {
                    u64 *xw = XW(bb, g, r);
                    st_agent(xw, ((u64)sq << 32) | (unsigned)(c >= 0 ? row0 + c : -1));
                    Tag<T>::put_s(xw + G, G, sq, cv);
                }
            }
        }
        if (g == 0 && wave == 0 && p != j) record_swap();
        // ---- 6. bulk rank-1 update of the other rows by waves 1-3, overlapping the
        //         exchange; wave 0 goes straight to the next column's poll.  The next
        //         column writes the other parity of prow2, and its first barrier
        //         orders this update before any use of the tile.
        //         (Single workgroup: no exchange to overlap, so wave 0 joins in.)
        if (solo) {
            if (j + 2 < kb) bulk_update<T>(tile, kbp, kb, lcol, prow, r1, Rg, j + 2, c, jn, t, NTK);
        } else if (wave > 0 && j + 2 < kb) {
            bulk_update<T>(tile, kbp, kb, lcol, prow, r1, Rg, j + 2, c, jn, t - 64, NTK - 64);
        }
    }
        }
    __syncthreads();

    for (int r = wave; r < Rg; r += NTK / 64)
        for (int c = lane; c < kb; c += 64) Ap[(long)(row0 + r) * a.lda + c] = tile[r * kbp + c];

    if (g == 0) {
        __syncthreads();
        if (wave == 0) {
            // the (destination, source) list of every moved row, compacted by wave 0
            // with ballots: the top rows first, then the rows outside them
            const int nout = s_nout, total = kb + nout;
            int np = 0;
            for (int base = 0; base < total; base += 64) {
                const int e = base + lane;
                int dst = 0, src = 0;
                bool moved = false;
                if (e < kb) {
                    dst = e; src = top_label[e]; moved = src != e;
                } else if (e < total) {
                    dst = out_pos[e - kb]; src = out_label[e - kb]; moved = src != dst;
                }
                const unsigned long long bl = __ballot(moved);
                const int at = np + __popcll(bl & ((1ull << lane) - 1ull));
                if (moved) { a.pairs[2 * at] = a.k0 + dst; a.pairs[2 * at + 1] = a.k0 + src; }
                np += __popcll(bl);
            }
            if (lane == 0) *a.npairs = np;
        }
    }
}

inline size_t lds_bytes(int R, int kb, size_t esz)
{
    return ((size_t)R * (kb + 1) + 3 * (size_t)kb + R) * esz + 3 * (size_t)kb * sizeof(int);
}

template <typename T>
inline int max_rows(int kb)
{
    const size_t budget = 160 * 1024 - 1024;   // static reduction buffers
    const size_t per_row = (size_t)(kb + 2) * sizeof(T);
    const size_t fixed = 3 * (size_t)kb * sizeof(T) + 3 * (size_t)kb * sizeof(int);
    return (int)((budget - fixed) / per_row);
}

} // namespace panel
