// Batched solves of many independent small systems of uneven size
// (n <= 256) in one launch.
//
// There is no batched solve in the reference (scirs2-linalg/src/batch/mod.rs
// offers matmul/matvec/add/sum only); each system follows the reference's
// per-system semantics: `solve` (scirs2-linalg/src/solve.rs:54-111) for the
// general systems and the cholesky + two triangular solves composition
// (decomposition.rs:44, solve.rs:139) for the SPD ones, with the same
// breakdown tests (|pivot| < eps, a_ii - sum <= 0).
//
// Arithmetic: the factor is computed in FP32 (blocked right-looking LU with
// partial pivoting, or lower Cholesky), then FP64 refinement — residual
// r = b - A x from the FP64 input, correction with the FP32 factors — until
// the normwise backward error of every right-hand side is <= tol.  A system
// whose FP32 factorisation breaks down, or whose refinement stalls, is
// factored again in FP64 by the same code (the breakdown is then decided in
// FP64, as the reference decides it).
//
// Scheduling: the host sorts the systems by decreasing n into a queue;
// workgroups pop queue entries with an atomic counter.  Entries with n >
// wave_max are processed by the whole 256-thread workgroup; smaller ones are
// processed one per wave (four per workgroup pop).  The crossover is a
// runtime argument picked by the batched autotuner (api side).
#include "common.hpp"
#include "kernels.hpp"
#include "prof.hpp"

namespace {

constexpr int NT = 256;
constexpr int NB = 32;          // panel width
constexpr int NMAX = 256;       // largest system handled

struct BatchDesc {
    int n, nrhs, spd, pad;
    long offA, offB;            // element offsets of A (n x n) and B/X (n x nrhs) in the packed buffers
    long offW;                  // element offset (FP64 units) of the factor workspace (n x n)
};

struct BatchArgs {
    const BatchDesc *desc;
    const int *order;           // queue: system indices sorted by decreasing n
    int batch;
    const double *A; const double *B; double *X;
    double *work;               // factor workspace, n*n doubles per system
    int *perm;                  // n ints per system (offset = offA / n ... use offB-like index: offP)
    const long *offP;
    int *status;                // per system: 0 ok, k>0 breakdown column/row, -5 not converged
    int *iters;                 // per system: refinement steps
    double *bwd;                // per system: final backward error
    unsigned *counter;          // queue head, reset to 0 before each launch
    int wave_max;               // n <= wave_max -> one wave per system
    int max_iter;
    double tol_scale;           // tol = tol_scale * n * 2^-53 (per system)
};

// ----------------------------------------------------------------- team ops
// A team is either the whole workgroup (TEAM = 256) or one wave (TEAM = 64).
template <int TEAM>
__device__ __forceinline__ void team_sync()
{
    if (TEAM == NT) __syncthreads();
    else __builtin_amdgcn_wave_barrier();
}

// arg-max of |v| over the team, lowest index on ties; result in all threads
template <int TEAM, typename T>
__device__ __forceinline__ void team_argmax(T &v, int &idx, T *red_v, int *red_i)
{
    wave_argmax(v, idx);
    if (TEAM == NT) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) { red_v[wave] = v; red_i[wave] = idx; }
        __syncthreads();
        v = red_v[0]; idx = red_i[0];
        for (int w = 1; w < NT / 64; ++w) {
            const int oi = red_i[w]; const T ov = red_v[w];
            if (oi >= 0 && (idx < 0 || ov > v || (ov == v && oi < idx))) { v = ov; idx = oi; }
        }
        __syncthreads();
    } else {
        v = __shfl(v, 0, 64);
        idx = __shfl(idx, 0, 64);
    }
}

template <int TEAM, typename T>
__device__ __forceinline__ T team_max(T v, T *red_v)
{
    v = wave_max(v);
    if (TEAM == NT) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) red_v[wave] = v;
        __syncthreads();
        v = fmax(fmax(red_v[0], red_v[1]), fmax(red_v[2], red_v[3]));
        __syncthreads();
    } else {
        v = __shfl(v, 0, 64);
    }
    return v;
}

// Blocked right-looking LU with partial pivoting of the n x n matrix M
// (row-major, ld n, global workspace).  The NB-column panel is factored in
// LDS (pan: n x (NB+1)), the trailing update reads the panel from LDS.
// Returns 0 or the 1-based breakdown column.
template <int TEAM, typename T>
__device__ int team_getrf(int n, T *M, int *perm, T *pan, T *red_v, int *red_i, int tid, T *ustage)
{
    for (int i = tid; i < n; i += TEAM) perm[i] = i;
    const T eps = num_limits<T>::eps();
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        const int mrows = n - k0;
        team_sync<TEAM>();
        // stage the panel rows k0..n, columns k0..k0+kb
        for (int idx = tid; idx < mrows * NB; idx += TEAM) {
            const int r = idx / NB, c = idx % NB;
            pan[r * (NB + 1) + c] = c < kb ? M[(long)(k0 + r) * n + k0 + c] : (T)0;
        }
        team_sync<TEAM>();
        for (int j = 0; j < kb; ++j) {
            T bv = (T)0; int bi = -1;
            for (int r = j + tid; r < mrows; r += TEAM) {
                const T v = fabs(pan[r * (NB + 1) + j]);
                if (bi < 0 || v > bv) { bv = v; bi = r; }
            }
            team_argmax<TEAM>(bv, bi, red_v, red_i);
            if (!(bv >= eps)) return k0 + j + 1;
            const int p = bi;
            if (p != j) {
                // swap inside the panel, and the whole row outside it
                for (int c = tid; c < NB; c += TEAM) {
                    const T a = pan[j * (NB + 1) + c];
                    pan[j * (NB + 1) + c] = pan[p * (NB + 1) + c];
                    pan[p * (NB + 1) + c] = a;
                }
                for (int c = tid; c < n; c += TEAM) {
                    if (c >= k0 && c < k0 + kb) continue;
                    const T a = M[(long)(k0 + j) * n + c];
                    M[(long)(k0 + j) * n + c] = M[(long)(k0 + p) * n + c];
                    M[(long)(k0 + p) * n + c] = a;
                }
                if (tid == 0) {
                    const int tmp = perm[k0 + j]; perm[k0 + j] = perm[k0 + p]; perm[k0 + p] = tmp;
                }
            }
            team_sync<TEAM>();
            const T piv = pan[j * (NB + 1) + j];
            for (int r = j + 1 + tid; r < mrows; r += TEAM) pan[r * (NB + 1) + j] /= piv;
            team_sync<TEAM>();
            const int ncol = kb - j - 1;
            if (ncol > 0)
                for (int w = tid; w < (mrows - j - 1) * ncol; w += TEAM) {
                    const int r = j + 1 + w / ncol, c = j + 1 + w % ncol;
                    pan[r * (NB + 1) + c] -= pan[r * (NB + 1) + j] * pan[j * (NB + 1) + c];
                }
            team_sync<TEAM>();
        }
        // write the panel back
        for (int idx = tid; idx < mrows * kb; idx += TEAM) {
            const int r = idx / kb, c = idx % kb;
            M[(long)(k0 + r) * n + k0 + c] = pan[r * (NB + 1) + c];
        }
        team_sync<TEAM>();
        const int c1 = k0 + kb;
        if (c1 >= n) continue;
        const int rows = n - c1, cols = n - c1;
        if constexpr (sizeof(T) == 4) {
        // U12 = L11^-1 A12 (rows k0..k0+kb): one column per thread, its kb
            // entries loaded together into registers, the same subtraction order
            // as the loop over i and q
            for (int c = c1 + tid; c < n; c += TEAM) {
                T u[NB];
#pragma unroll
                for (int i = 0; i < NB; ++i) u[i] = i < kb ? M[(long)(k0 + i) * n + c] : (T)0;
#pragma unroll
                for (int i = 1; i < NB; ++i) {
                    // row i of L11 through an offset the compiler sees only once
                    // u[i - 1] exists, so its loads are not all hoisted (they spill)
                    int off = i * (NB + 1);
                    asm volatile("" : "+v"(off) : "v"(u[i - 1]));
                    const T *li = pan + off;
                    T sv = u[i];
#pragma unroll
                    for (int q = 0; q < i; ++q) sv -= li[q] * u[q];
                    u[i] = sv;
                }
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    if (i < kb) M[(long)(k0 + i) * n + c] = u[i];
                    // rows past kb are zero in the staged copy (the panel's columns past kb are zero too)
                    if (ustage) ustage[i * cols + (c - c1)] = i < kb ? u[i] : (T)0;
                }
            }
        } else {
            // the FP64 fallback: the column in memory (registers for 32 doubles
            // would spill in this kernel)
            for (int c = c1 + tid; c < n; c += TEAM)
                for (int i = 1; i < kb; ++i) {
                    T sv = M[(long)(k0 + i) * n + c];
                    for (int q = 0; q < i; ++q) sv -= pan[i * (NB + 1) + q] * M[(long)(k0 + q) * n + c];
                    M[(long)(k0 + i) * n + c] = sv;
                }
        }
        team_sync<TEAM>();
        // A22 -= L21 U12 (U12 from the staged copy when there is one)
        for (int w = tid; w < rows * cols; w += TEAM) {
            const int r = w / cols, c = w % cols;
            const T *lr = pan + (kb + r) * (NB + 1);
            T sv = (T)0;
            if (ustage) {
#pragma unroll 8
                for (int q = 0; q < NB; ++q) sv += lr[q] * ustage[q * cols + c];
            } else {
#pragma unroll 8
                for (int q = 0; q < NB; ++q)
                    if (q < kb) sv += lr[q] * M[(long)(k0 + q) * n + c1 + c];
            }
            M[(long)(c1 + r) * n + c1 + c] -= sv;
        }
    }
    team_sync<TEAM>();
    return 0;
}

// Lower Cholesky (right-looking, unblocked over NB-column panels in LDS), only
// the lower triangle of M is read.  Returns 0 or the 1-based failing row.
template <int TEAM, typename T>
__device__ int team_potrf(int n, T *M, T *pan, int tid)
{
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        const int mrows = n - k0;
        team_sync<TEAM>();
        for (int idx = tid; idx < mrows * NB; idx += TEAM) {
            const int r = idx / NB, c = idx % NB;
            pan[r * (NB + 1) + c] = (c < kb && c <= r) ? M[(long)(k0 + r) * n + k0 + c] : (T)0;
        }
        team_sync<TEAM>();
        for (int j = 0; j < kb; ++j) {
            const T d = pan[j * (NB + 1) + j];
            if (!(d > (T)0)) return k0 + j + 1;
            const T l = sqrt(d);
            team_sync<TEAM>();
            for (int r = j + tid; r < mrows; r += TEAM) pan[r * (NB + 1) + j] = (r == j) ? l : pan[r * (NB + 1) + j] / l;
            team_sync<TEAM>();
            const int ncol = kb - j - 1;
            if (ncol > 0)
                for (int w = tid; w < (mrows - j - 1) * ncol; w += TEAM) {
                    const int r = j + 1 + w / ncol, c = j + 1 + w % ncol;
                    if (c <= r) pan[r * (NB + 1) + c] -= pan[r * (NB + 1) + j] * pan[c * (NB + 1) + j];
                }
            team_sync<TEAM>();
        }
        for (int idx = tid; idx < mrows * kb; idx += TEAM) {
            const int r = idx / kb, c = idx % kb;
            if (c <= r) M[(long)(k0 + r) * n + k0 + c] = pan[r * (NB + 1) + c];
        }
        team_sync<TEAM>();
        const int c1 = k0 + kb;
        if (c1 >= n) continue;
        // A22 (lower) -= L21 L21^T
        const int rows = n - c1;
        for (int w = tid; w < rows * rows; w += TEAM) {
            const int r = w / rows, c = w % rows;
            if (c > r) continue;
            const T *lr = pan + (kb + r) * (NB + 1);
            const T *lc = pan + (kb + c) * (NB + 1);
            T sv = (T)0;
            // the panel's columns past kb are zero
#pragma unroll 8
            for (int q = 0; q < NB; ++q) sv += lr[q] * lc[q];
            M[(long)(c1 + r) * n + c1 + c] -= sv;
        }
    }
    team_sync<TEAM>();
    return 0;
}

__device__ __forceinline__ double bcast_lane(double v, int q)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), q);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), q);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// d <- op^-1 d with the factors in M (row-major, ld n); d in LDS (length n).
// LU: L unit lower then U; Cholesky: L then L^T.  Blocks of NB rows: the
// first wave solves the diagonal block (lane i = row i, the solved value
// broadcast by readlane), then every thread takes one row outside the block
// and removes the block's contribution.  Each row receives the same
// subtractions, in the same order, as a column-by-column sweep.
template <int TEAM, typename T, typename V>
__device__ void team_subst(int n, const T *M, V *d, bool chol, int tid)
{
    const int lane = tid & 63;
    const bool w0 = tid < 64;
    // forward: L (unit for LU)
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        if (w0) {
            const int i = lane;
            const bool in = i < kb;
            const int ri = k0 + min(i, kb - 1);
            V di = in ? d[k0 + i] : (V)0;
            T lrow[NB];
#pragma unroll
            for (int q = 0; q < NB; ++q) lrow[q] = M[(long)ri * n + k0 + min(q, kb - 1)];
            const V dg = (V)M[(long)ri * n + ri];
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                if (q < kb) {
                    if (chol && i == q) di /= dg;
                    const V xq = bcast_lane(di, q);
                    if (in && i > q) di -= (V)lrow[q] * xq;
                }
            }
            if (in) d[k0 + i] = di;
        }
        team_sync<TEAM>();
        for (int r = k0 + kb + tid; r < n; r += TEAM) {
            V sv = d[r];
#pragma unroll 8
            for (int q = 0; q < NB; ++q)
                if (q < kb) sv -= (V)M[(long)r * n + k0 + q] * d[k0 + q];
            d[r] = sv;
        }
        team_sync<TEAM>();
    }
    // backward: U (LU) or L^T (Cholesky), blocks from the bottom
    for (int k0 = (n - 1) / NB * NB; k0 >= 0; k0 -= NB) {
        const int kb = min(NB, n - k0);
        if (w0) {
            const int i = lane;
            const bool in = i < kb;
            const int ri = k0 + min(i, kb - 1);
            V di = in ? d[k0 + i] : (V)0;
            T urow[NB];
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const int cq = k0 + min(q, kb - 1);
                urow[q] = chol ? M[(long)cq * n + ri] : M[(long)ri * n + cq];
            }
            const V dg = (V)M[(long)ri * n + ri];
#pragma unroll
            for (int q = NB - 1; q >= 0; --q) {
                if (q < kb) {
                    if (i == q) di /= dg;
                    const V xq = bcast_lane(di, q);
                    if (in && i < q) di -= (V)urow[q] * xq;
                }
            }
            if (in) d[k0 + i] = di;
        }
        team_sync<TEAM>();
        for (int r = tid; r < k0; r += TEAM) {
            V sv = d[r];
#pragma unroll 8
            for (int q = NB - 1; q >= 0; --q)
                if (q < kb) sv -= (V)(chol ? M[(long)(k0 + q) * n + r] : M[(long)r * n + k0 + q]) * d[k0 + q];
            d[r] = sv;
        }
        team_sync<TEAM>();
    }
}

// The FP64 fallback's solve, column by column (d in LDS, length n): every thread reads the finished d[i], then the team
// removes its contribution from the remaining entries (one sync per row).
// LU: L unit lower then U; Cholesky: L then L^T.
template <int TEAM, typename T, typename V>
__device__ void team_subst_cols(int n, const T *M, V *d, bool chol, int tid)
{
    for (int i = 0; i < n; ++i) {
        V xi = d[i];
        if (chol) xi /= (V)M[(long)i * n + i];
        for (int r = i + 1 + tid; r < n; r += TEAM) d[r] -= (V)M[(long)r * n + i] * xi;
        team_sync<TEAM>();
        if (tid == 0 && chol) d[i] = xi;
        team_sync<TEAM>();
    }
    for (int i = n - 1; i >= 0; --i) {
        const V xi = d[i] / (V)M[(long)i * n + i];
        for (int r = tid; r < i; r += TEAM)
            d[r] -= (V)(chol ? M[(long)i * n + r] : M[(long)r * n + i]) * xi;
        team_sync<TEAM>();
        if (tid == 0) d[i] = xi;
        team_sync<TEAM>();
    }
}

// Solve one system with a team: FP32 factor + FP64 refinement, FP64 fallback.
template <int TEAM>
__device__ void team_solve(const BatchArgs &a, int sys, float *pan32, double *pan64, double *vec, double *red,
                           float *redf, int *redi, int tid)
{
    const BatchDesc d = a.desc[sys];
    const int n = d.n, nrhs = d.nrhs;
    const double *A = a.A + d.offA;
    const double *B = a.B + d.offB;
    double *X = a.X + d.offB;
    int *perm = a.perm + a.offP[sys];
    double *W64 = a.work + d.offW;
    float *W32 = reinterpret_cast<float *>(W64);
    const double tol = a.tol_scale * n * 0x1p-53;

    // ||A||_inf
    double an = 0.0;
    for (int r = tid; r < n; r += TEAM) {
        double s = 0.0;
        for (int c = 0; c < n; ++c) s += fabs(A[(long)r * n + c]);
        an = fmax(an, s);
    }
    an = team_max<TEAM>(an, red);

    int status = 0, it_used = 0;
    double worst = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        const bool f64 = (pass == 1);
        bool chol = d.spd != 0;
        // copy A into the workspace and factor
        int rc;
        if (!f64) {
            for (int idx = tid; idx < n * n; idx += TEAM) W32[idx] = (float)A[idx];
            team_sync<TEAM>();
            // U12 is staged in the half of the panel buffer an FP32 panel leaves free
            rc = chol ? team_potrf<TEAM, float>(n, W32, pan32, tid)
                      : team_getrf<TEAM, float>(n, W32, perm, pan32, redf, redi, tid,
                                                pan32 + (size_t)(TEAM == NT ? NMAX : NMAX / 4) * (NB + 1));
            if (rc != 0) continue;   // FP32 breakdown: the same factorisation decides in FP64
        } else {
            for (int idx = tid; idx < n * n; idx += TEAM) W64[idx] = A[idx];
            team_sync<TEAM>();
            if (chol) {
                rc = team_potrf<TEAM, double>(n, W64, pan64, tid);
            } else {
                rc = team_getrf<TEAM, double>(n, W64, perm, pan64, red, redi, tid, nullptr);
            }
            if (rc != 0) { status = rc; break; }
        }
        // refinement for each right-hand side
        worst = 0.0;
        bool all_ok = true;
        for (int j = 0; j < nrhs; ++j) {
            for (int r = tid; r < n; r += TEAM) X[(long)r * nrhs + j] = 0.0;
            team_sync<TEAM>();
            double bn = 0.0;
            for (int r = tid; r < n; r += TEAM) bn = fmax(bn, fabs(B[(long)r * nrhs + j]));
            bn = team_max<TEAM>(bn, red);
            double eta = 1.0, prev = INFINITY;
            int it = 0;
            for (; it <= a.max_iter; ++it) {
                // r = b - A x  (FP64), stored in vec: one wave per row, the
                // lanes over the columns (coalesced row reads)
                double rn = 0.0, xn = 0.0;
                {
                    const int lane = tid & 63;
                    for (int r = tid >> 6; r < n; r += TEAM / 64) {
                        double ax = 0.0;
                        for (int c = lane; c < n; c += 64) ax += A[(long)r * n + c] * X[(long)c * nrhs + j];
                        ax = wave_sum(ax);
                        const double rv = B[(long)r * nrhs + j] - ax;
                        if (lane == 0) vec[r] = rv;
                        rn = fmax(rn, fabs(rv));
                    }
                    for (int r = tid; r < n; r += TEAM) xn = fmax(xn, fabs(X[(long)r * nrhs + j]));
                }
                team_sync<TEAM>();
                rn = team_max<TEAM>(rn, red);
                xn = team_max<TEAM>(xn, red);
                const double den = an * xn + bn;
                eta = den > 0 ? rn / den : (rn > 0 ? INFINITY : 0.0);
                if (eta <= tol) break;
                if (it == a.max_iter || (it >= 2 && !(eta < 0.5 * prev))) break;
                prev = eta;
                // permute for LU: d = P r
                if (!chol) {
                    team_sync<TEAM>();
                    double *tmp = vec + NMAX;
                    for (int r = tid; r < n; r += TEAM) tmp[r] = vec[perm[r]];
                    team_sync<TEAM>();
                    for (int r = tid; r < n; r += TEAM) vec[r] = tmp[r];
                }
                team_sync<TEAM>();
                if (f64) team_subst_cols<TEAM, double, double>(n, W64, vec, chol, tid);
                else team_subst<TEAM, float, double>(n, W32, vec, chol, tid);
                for (int r = tid; r < n; r += TEAM) X[(long)r * nrhs + j] += vec[r];
                team_sync<TEAM>();
            }
            it_used = max(it_used, it);
            worst = fmax(worst, eta);
            if (!(eta <= tol)) all_ok = false;
        }
        if (all_ok || f64) {
            status = all_ok ? 0 : (f64 ? 0 : -5);
            break;
        }
    }
    if (tid == 0) {
        a.status[sys] = status;
        a.iters[sys] = it_used;
        a.bwd[sys] = worst;
    }
    team_sync<TEAM>();
}

__global__ void batched_reset_kernel(unsigned *counter) { *counter = 0u; }

__global__ __launch_bounds__(NT) void batched_solve_kernel(BatchArgs a)
{
    // one NMAX x (NB+1) FP64 panel for a whole-workgroup system, or four
    // (NMAX/4) x (NB+1) slices, one per wave, for small systems
    __shared__ double panbuf[NMAX * (NB + 1)];
    __shared__ double vec[4][2 * NMAX];
    __shared__ double red[8];
    __shared__ float redf[8];
    __shared__ int redi[8];
    __shared__ int s_item;
    constexpr int SLICE = NMAX / 4 * (NB + 1);   // doubles per wave slice
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    while (true) {
        if (t == 0) s_item = (int)atomicAdd(a.counter, 4u);
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= a.batch) break;
        for (int q = 0; q < 4 && item + q < a.batch; ++q) {
            const int sys = a.order[item + q];
            if (a.desc[sys].n > a.wave_max) {
                team_solve<NT>(a, sys, reinterpret_cast<float *>(panbuf), panbuf, &vec[0][0], red, redf, redi, t);
            } else {
                // the remaining small entries of this pop: one per wave
                const int qq = q + wave;
                if (qq < 4 && item + qq < a.batch) {
                    const int s2 = a.order[item + qq];
                    double *slice = panbuf + (size_t)wave * SLICE;
                    team_solve<64>(a, s2, reinterpret_cast<float *>(slice), slice, vec[wave], red + 2 * wave,
                                   redf + 2 * wave, redi + 2 * wave, lane);
                }
                __syncthreads();
                break;
            }
        }
        __syncthreads();
    }
}
} // namespace

// ===================================================================== host
#include "api_internal.hpp"
#include <string>
int mkdir_p(const std::string &dir);
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

struct BatchPlan {
    uint64_t key = 0;
    int batch = 0;
    BatchDesc *d_desc = nullptr;
    int *d_order = nullptr;
    long *d_offP = nullptr;
    int *d_status = nullptr, *d_iters = nullptr;
    double *d_bwd = nullptr;
    double *d_work = nullptr;
    int *d_perm = nullptr;
    unsigned *d_counter = nullptr;
    size_t cap_batch = 0, cap_work = 0, cap_perm = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_wave_max = -1;
    std::string tune_key;
};

BatchPlan g_plan;
std::map<std::string, int> g_tune;      // size-histogram key -> wave_max
bool g_tune_loaded = false;

std::string tune_path()
{
    const char *p = getenv("SCIRS_GPU_TUNE_FILE");
    if (p && *p) return p;
    const char *h = getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.cache/scirs_gpu_batched_tune.json";
}

void tune_load()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    FILE *f = fopen(tune_path().c_str(), "r");
    if (!f) return;
    char key[256];
    int v;
    // the file is a flat JSON object {"key": value, ...}; read it token by token
    int c;
    while ((c = fgetc(f)) != EOF) {
        if (c != '"') continue;
        int k = 0;
        while ((c = fgetc(f)) != EOF && c != '"' && k < 255) key[k++] = (char)c;
        key[k] = 0;
        while ((c = fgetc(f)) != EOF && c != ':') {}
        if (fscanf(f, "%d", &v) == 1) g_tune[key] = v;
    }
    fclose(f);
}

void tune_save()
{
    std::string path = tune_path();
    std::string dir = path.substr(0, path.find_last_of('/'));
    if (!dir.empty()) {
        (void)mkdir_p(dir);   // create the cache directory if needed
    }
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return;
    fprintf(f, "{");
    bool first = true;
    for (auto &kv : g_tune) {
        fprintf(f, "%s\"%s\": %d", first ? "" : ", ", kv.first.c_str(), kv.second);
        first = false;
    }
    fprintf(f, "}\n");
    fclose(f);
}

uint64_t fnv1a(uint64_t h, const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

template <typename T>
int grow(T **p, size_t *cap, size_t count)
{
    if (*cap >= count && *p) return SCIRS_OK;
    if (*p) {
        SCIRS_HIP_CHECK(hipDeviceSynchronize());
        SCIRS_HIP_CHECK(hipFree(*p));
        *p = nullptr;
    }
    if (hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
        scirs_set_error("batched: hipMalloc of %zu bytes failed", count * sizeof(T));
        *cap = 0;
        return SCIRS_ERR_NOMEM;
    }
    *cap = count;
    return SCIRS_OK;
}

int launch_batch(BatchPlan &pl, const double *dA, const double *dB, double *dX, int wave_max, int max_iter,
                 double tol_scale, hipStream_t s)
{
    BatchArgs a;
    a.desc = pl.d_desc; a.order = pl.d_order; a.batch = pl.batch;
    a.A = dA; a.B = dB; a.X = dX; a.work = pl.d_work; a.perm = pl.d_perm; a.offP = pl.d_offP;
    a.status = pl.d_status; a.iters = pl.d_iters; a.bwd = pl.d_bwd; a.counter = pl.d_counter;
    a.wave_max = wave_max; a.max_iter = max_iter; a.tol_scale = tol_scale;
    hipLaunchKernelGGL(batched_reset_kernel, dim3(1), dim3(1), 0, s, pl.d_counter);
    // a persistent grid: a few workgroups per CU pull work from the queue
    const int grid = std::min(pl.batch, scirs::ctx().cu_count * 2);
    hipLaunchKernelGGL(batched_solve_kernel, dim3(std::max(grid, 1)), dim3(NT), 0, s, a);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

} // namespace

static int mkdir_p_impl(const std::string &dir);
int mkdir_p(const std::string &dir) { return mkdir_p_impl(dir); }
#include <sys/stat.h>
static int mkdir_p_impl(const std::string &dir)
{
    std::string cur;
    for (size_t i = 0; i < dir.size(); ++i) {
        cur += dir[i];
        if (dir[i] == '/' || i + 1 == dir.size()) mkdir(cur.c_str(), 0755);
    }
    return 0;
}

extern "C" int scirs_gpu_dgesv_batched_dev(int batch, const int *n, const int *nrhs, const unsigned char *spd,
                                           const double *dA, const int64_t *offA, const double *dB,
                                           const int64_t *offB, double *dX, int *status, double *bwd,
                                           int max_iter, double tol_scale, void *stream, scirs_batch_info *info)
{
    using namespace scirs;
    const double t0 = now_ms();
    if (info) memset(info, 0, sizeof(*info));
    if (batch <= 0 || !n || !nrhs || !dA || !offA || !dB || !offB || !dX || !status) {
        scirs_set_error("dgesv_batched: invalid arguments");
        return SCIRS_ERR_ARG;
    }
    for (int i = 0; i < batch; ++i)
        if (n[i] < 1 || n[i] > NMAX || nrhs[i] < 1 || nrhs[i] > 16) {
            scirs_set_error("dgesv_batched: system %d has n=%d nrhs=%d (1 <= n <= %d, 1 <= nrhs <= 16)", i, n[i],
                            nrhs[i], NMAX);
            return SCIRS_ERR_ARG;
        }
    if (max_iter <= 0) max_iter = 10;
    if (!(tol_scale > 0)) tol_scale = 1.0;
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    Context &c = ctx();
    int rc = c.init(0);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;   // NULL: the default stream, ordered after the caller's work on it
    BatchPlan &pl = g_plan;

    // ---- plan: descriptors, size-sorted queue, workspace (rebuilt when the batch changes)
    uint64_t key = 1469598103934665603ull;
    key = fnv1a(key, &batch, sizeof(batch));
    key = fnv1a(key, n, sizeof(int) * batch);
    key = fnv1a(key, nrhs, sizeof(int) * batch);
    if (spd) key = fnv1a(key, spd, batch);
    key = fnv1a(key, offA, sizeof(int64_t) * batch);
    key = fnv1a(key, offB, sizeof(int64_t) * batch);
    const void *ptrs[3] = {dA, dB, dX};
    key = fnv1a(key, ptrs, sizeof(ptrs));
    key = fnv1a(key, &max_iter, sizeof(max_iter));
    key = fnv1a(key, &tol_scale, sizeof(tol_scale));
    const bool same = (pl.key == key && pl.batch == batch);
    if (!same) {
        std::vector<BatchDesc> desc(batch);
        std::vector<long> offP(batch);
        long work = 0, pw = 0;
        long nsum = 0, nmax = 0;
        for (int i = 0; i < batch; ++i) {
            desc[i].n = n[i]; desc[i].nrhs = nrhs[i]; desc[i].spd = spd ? spd[i] : 0; desc[i].pad = 0;
            desc[i].offA = offA[i]; desc[i].offB = offB[i]; desc[i].offW = work;
            work += (long)n[i] * n[i];
            offP[i] = pw; pw += n[i];
            nsum += n[i]; nmax = std::max<long>(nmax, n[i]);
        }
        std::vector<int> order(batch);
        for (int i = 0; i < batch; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
        if ((rc = grow(&pl.d_desc, &pl.cap_batch, batch))) return rc;
        size_t cap2 = 0, cap3 = 0, cap4 = 0, cap5 = 0, cap6 = 0;
        // the per-system arrays are (re)allocated together with the descriptors
        if (pl.d_order) { (void)hipFree(pl.d_order); (void)hipFree(pl.d_offP); (void)hipFree(pl.d_status);
                          (void)hipFree(pl.d_iters); (void)hipFree(pl.d_bwd); pl.d_order = nullptr; pl.d_offP = nullptr;
                          pl.d_status = nullptr; pl.d_iters = nullptr; pl.d_bwd = nullptr; }
        if ((rc = grow(&pl.d_order, &cap2, batch)) || (rc = grow(&pl.d_offP, &cap3, batch)) ||
            (rc = grow(&pl.d_status, &cap4, batch)) || (rc = grow(&pl.d_iters, &cap5, batch)) ||
            (rc = grow(&pl.d_bwd, &cap6, batch)))
            return rc;
        if ((rc = grow(&pl.d_work, &pl.cap_work, (size_t)work))) return rc;
        if ((rc = grow(&pl.d_perm, &pl.cap_perm, (size_t)pw))) return rc;
        if (!pl.d_counter) {
            size_t cc = 0;
            if ((rc = grow(&pl.d_counter, &cc, 64))) return rc;
        }
        SCIRS_HIP_CHECK(hipMemcpy(pl.d_desc, desc.data(), sizeof(BatchDesc) * batch, hipMemcpyHostToDevice));
        SCIRS_HIP_CHECK(hipMemcpy(pl.d_order, order.data(), sizeof(int) * batch, hipMemcpyHostToDevice));
        SCIRS_HIP_CHECK(hipMemcpy(pl.d_offP, offP.data(), sizeof(long) * batch, hipMemcpyHostToDevice));
        if (pl.exec) { (void)hipGraphExecDestroy(pl.exec); pl.exec = nullptr; }
        if (pl.graph) { (void)hipGraphDestroy(pl.graph); pl.graph = nullptr; }
        pl.key = key;
        pl.batch = batch;
        char tk[128];
        snprintf(tk, sizeof(tk), "gfx950:nmax%ld:nmean%ld", (nmax + 31) / 32 * 32, (nsum / batch + 31) / 32 * 32);
        pl.tune_key = tk;
    }

    // ---- schedule crossover: explicit, tuned table, or tune now
    int wave_max = -1;
    const char *wm = getenv("SCIRS_GPU_BATCH_WAVE_MAX");
    if (wm && *wm) wave_max = atoi(wm);
    if (wave_max < 0) {
        tune_load();
        auto it = g_tune.find(pl.tune_key);
        if (it != g_tune.end()) wave_max = it->second;
    }
    if (wave_max < 0) {
        const int cand[3] = {0, 32, 64};
        float best = 1e30f;
        scirs::CallEvents ev;
        if (int erc = ev.create(2)) return erc;
        hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
        for (int ci = 0; ci < 3; ++ci) {
            SCIRS_HIP_CHECK(hipEventRecord(e0, s));
            if ((rc = launch_batch(pl, dA, dB, dX, cand[ci], max_iter, tol_scale, s))) return rc;
            SCIRS_HIP_CHECK(hipEventRecord(e1, s));
            SCIRS_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            SCIRS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; wave_max = cand[ci]; }
        }
        g_tune[pl.tune_key] = wave_max;
        tune_save();
    }

    // a wave's LDS slice holds systems up to NMAX / 4 rows
    wave_max = std::max(0, std::min(wave_max, NMAX / 4));

    // ---- one graph per plan and crossover: [reset queue, persistent solve], replayed
    bool replay = true;
    if (!pl.exec || pl.graph_wave_max != wave_max) {
        replay = false;
        if (pl.exec) { (void)hipGraphExecDestroy(pl.exec); pl.exec = nullptr; }
        if (pl.graph) { (void)hipGraphDestroy(pl.graph); pl.graph = nullptr; }
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
        SCIRS_HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
        rc = launch_batch(pl, dA, dB, dX, wave_max, max_iter, tol_scale, c.stream);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(c.stream, &g);
        if (rc) return rc;
        if (ce != hipSuccess) {
            scirs_set_error("batched: graph capture failed: %s", hipGetErrorString(ce));
            return SCIRS_ERR_HIP;
        }
        pl.graph = g;
        SCIRS_HIP_CHECK(hipGraphInstantiate(&pl.exec, pl.graph, nullptr, nullptr, 0));
        pl.graph_wave_max = wave_max;
    }
    scirs::CallEvents ev;
    if (int erc = ev.create(2)) return erc;
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    SCIRS_HIP_CHECK(hipEventRecord(e0, s));
    {
        scirs::ProfScope prof(SCIRS_PROF_OTHER, 0.0, 0.0, s);
        SCIRS_HIP_CHECK(hipGraphLaunch(pl.exec, s));
    }
    SCIRS_HIP_CHECK(hipEventRecord(e1, s));
    SCIRS_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    SCIRS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));

    std::vector<int> iters(batch);
    std::vector<double> bw(batch);
    SCIRS_HIP_CHECK(hipMemcpy(status, pl.d_status, sizeof(int) * batch, hipMemcpyDeviceToHost));
    SCIRS_HIP_CHECK(hipMemcpy(iters.data(), pl.d_iters, sizeof(int) * batch, hipMemcpyDeviceToHost));
    SCIRS_HIP_CHECK(hipMemcpy(bw.data(), pl.d_bwd, sizeof(double) * batch, hipMemcpyDeviceToHost));
    if (bwd) memcpy(bwd, bw.data(), sizeof(double) * batch);
    int failed = 0, maxit = 0;
    double maxb = 0;
    for (int i = 0; i < batch; ++i) {
        failed += status[i] != 0;
        maxit = std::max(maxit, iters[i]);
        if (status[i] == 0) maxb = std::max(maxb, bw[i]);
    }
    if (info) {
        info->failed = failed;
        info->max_iters = maxit;
        info->max_backward_error = maxb;
        info->time_ms = ms;
        info->wave_max = wave_max;
        info->graph_replayed = replay ? 1 : 0;
        info->total_ms = now_ms() - t0;
    }
    return SCIRS_OK;
}

// ---------------------------------------------------------------- host memory
// Host-memory batched solve: the systems are packed into pinned staging, sent
// in one copy, solved by the device entry point above, and the solutions come
// back in one copy.  Staging (pinned and device) is kept between calls and
// grown on demand, so a repeated batch of the same shape replays the cached
// graph on the same device buffers.
namespace {
struct HostBatch {
    double *h = nullptr;   // pinned: packed A | packed B (X on the way back)
    size_t hcap = 0;
    double *d = nullptr;   // device: packed A | packed B | packed X
    size_t dcap = 0;
};
HostBatch g_hb;

int grow_pinned(double **p, size_t *cap, size_t count)
{
    if (*cap >= count && *p) return SCIRS_OK;
    if (*p) {
        SCIRS_HIP_CHECK(hipDeviceSynchronize());
        SCIRS_HIP_CHECK(hipHostFree(*p));
        *p = nullptr;
    }
    if (hipHostMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        scirs_set_error("batched: pinned allocation of %zu bytes failed", count * sizeof(double));
        *cap = 0;
        return SCIRS_ERR_NOMEM;
    }
    *cap = count;
    return SCIRS_OK;
}
} // namespace

extern "C" int scirs_gpu_dgesv_batched(int batch, const int *n, const int *nrhs, const unsigned char *spd,
                                       const double *const *A, const int64_t *lda, const double *const *B,
                                       const int64_t *ldb, double *const *X, const int64_t *ldx, int *status,
                                       double *bwd, int max_iter, double tol_scale, scirs_batch_info *info)
{
    using namespace scirs;
    const double t0 = now_ms();
    if (info) memset(info, 0, sizeof(*info));
    if (batch <= 0 || !n || !nrhs || !A || !lda || !B || !ldb || !X || !ldx || !status) {
        scirs_set_error("dgesv_batched: invalid arguments");
        return SCIRS_ERR_ARG;
    }
    std::vector<int64_t> offA(batch), offB(batch);
    size_t na = 0, nb = 0;
    for (int i = 0; i < batch; ++i) {
        if (n[i] < 1 || n[i] > NMAX || nrhs[i] < 1 || nrhs[i] > 16 || !A[i] || !B[i] || !X[i] || lda[i] < n[i] ||
            ldb[i] < nrhs[i] || ldx[i] < nrhs[i]) {
            scirs_set_error("dgesv_batched: system %d has n=%d nrhs=%d lda=%lld ldb=%lld ldx=%lld or a null pointer "
                            "(1 <= n <= %d, 1 <= nrhs <= 16, leading dimensions >= the row length)",
                            i, n[i], nrhs[i], (long long)lda[i], (long long)ldb[i], (long long)ldx[i], NMAX);
            return SCIRS_ERR_ARG;
        }
        offA[i] = (int64_t)na;
        offB[i] = (int64_t)nb;
        na += (size_t)n[i] * n[i];
        nb += (size_t)n[i] * nrhs[i];
    }
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    Context &c = ctx();
    if (int rc = c.init(0)) return rc;
    hipStream_t s = c.stream;
    HostBatch &hb = g_hb;
    if (int rc = grow_pinned(&hb.h, &hb.hcap, na + nb)) return rc;
    if (int rc = grow(&hb.d, &hb.dcap, na + 2 * nb)) return rc;
    double *hA = hb.h, *hB = hb.h + na;
    for (int i = 0; i < batch; ++i) {
        const int ni = n[i], ri = nrhs[i];
        for (int r = 0; r < ni; ++r) {
            memcpy(hA + offA[i] + (size_t)r * ni, A[i] + (size_t)r * lda[i], sizeof(double) * ni);
            memcpy(hB + offB[i] + (size_t)r * ri, B[i] + (size_t)r * ldb[i], sizeof(double) * ri);
        }
    }
    double *dA = hb.d, *dB = hb.d + na, *dX = hb.d + na + nb;
    SCIRS_HIP_CHECK(hipMemcpyAsync(dA, hA, sizeof(double) * (na + nb), hipMemcpyHostToDevice, s));
    int rc = scirs_gpu_dgesv_batched_dev(batch, n, nrhs, spd, dA, offA.data(), dB, offB.data(), dX, status, bwd,
                                         max_iter, tol_scale, s, info);
    if (rc) return rc;
    SCIRS_HIP_CHECK(hipMemcpyAsync(hB, dX, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < batch; ++i) {
        const int ni = n[i], ri = nrhs[i];
        for (int r = 0; r < ni; ++r)
            memcpy(X[i] + (size_t)r * ldx[i], hB + offB[i] + (size_t)r * ri, sizeof(double) * ri);
    }
    if (info) info->total_ms = now_ms() - t0;
    return SCIRS_OK;
}
