// Schur-complement update on the gfx950 matrix cores, FP64 and FP32:
//     C[M x N] -= A[M x K] * op(B)      (all row-major)
// op(B) = B (K x N)  or  B^T with B stored N x K (the Cholesky SYRK case).
//
// Tiling: 128 x 128 output tile per 256-thread workgroup (4 wave64s in a 2 x 2
// arrangement, 64 x 64 per wave = 4 x 4 v_mfma_{f64,f32}_16x16x4 blocks), K
// staged through LDS in slabs of 16 with a register prefetch of the next slab,
// so each slab costs one barrier.  Both operand tiles live k-major in LDS,
// which makes every MFMA operand read a run of 16 consecutive elements.
//
// Workgroups are renumbered so that the tiles one XCD works on share A rows
// and B columns in that XCD's L2: hardware dispatches workgroup w to XCD w % 8.
#include "common.hpp"
#include "kernels.hpp"
#include "tile_map.hpp"
#include <cstdlib>

namespace {

constexpr int BM = 128, BN = 128, BK = 16;
constexpr int PAD = 4;
constexpr int LDA_S = BM + PAD;
constexpr int LDB_S = BN + PAD;
constexpr int NTHREADS = 256;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct acc_t;
template <> struct acc_t<double> { typedef d4 type; };
template <> struct acc_t<float> { typedef f4 type; };

__device__ __forceinline__ d4 mfma16x16x4(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma16x16x4(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T>
struct GemmArgs {
    int M, N, K;
    const T *A; long lda;
    const T *B; long ldb;
    T *C; long ldc;
    int tiles_m, tiles_n;
    int lower;
    const int *info;
};

// CPRE: prefetch C before the K loop (one workgroup per CU: the prefetch and
// the accumulators need the 512-register budget of one wave per SIMD).
template <typename T, bool TRANS_B, bool CPRE>
__global__ __launch_bounds__(NTHREADS, CPRE ? 1 : 2) void gemm_sub_kernel(GemmArgs<T> g)
{
    if (g.info && *g.info != 0) return;
    int tm, tn;
    scirs_tile_of(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n, tm, tn);
    if (tm >= g.tiles_m || tn >= g.tiles_n) return;
    const int row0 = tm * BM, col0 = tn * BN;
    if (g.lower && col0 > row0 + BM - 1) return;
    typedef typename acc_t<T>::type V4;

    __shared__ T As[2][BK * LDA_S];
    __shared__ T Bs[2][BK * LDB_S];

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    const int a_r = t >> 1, a_k = (t & 1) * 8;
    const int b_k = t >> 4, b_c = (t & 15) * 8;
    T ra[8], rb[8];

    auto load_slab = [&](int k0) {
        {
            int gr = row0 + a_r;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                int gk = k0 + a_k + i;
                ra[i] = (gr < g.M && gk < g.K) ? g.A[(long)gr * g.lda + gk] : (T)0;
            }
        }
        if (!TRANS_B) {
            int gk = k0 + b_k;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                int gc = col0 + b_c + i;
                rb[i] = (gk < g.K && gc < g.N) ? g.B[(long)gk * g.ldb + gc] : (T)0;
            }
        } else {
            int gc = col0 + a_r;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                int gk = k0 + a_k + i;
                rb[i] = (gc < g.N && gk < g.K) ? g.B[(long)gc * g.ldb + gk] : (T)0;
            }
        }
    };
    auto store_slab = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 8; ++i) As[buf][(a_k + i) * LDA_S + a_r] = ra[i];
        if (!TRANS_B) {
#pragma unroll
            for (int i = 0; i < 8; ++i) Bs[buf][b_k * LDB_S + b_c + i] = rb[i];
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) Bs[buf][(a_k + i) * LDB_S + a_r] = rb[i];
        }
    };

    V4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = V4{0, 0, 0, 0};

    const int lr = lane & 15;
    const int lk = lane >> 4;
    // C is read once: its loads are issued before the K loop so their latency
    // overlaps the matrix-core work.  Lane l, register q of a 16x16 block holds
    // column l%16 and row 4*(l/16)+q for v_mfma_f32_16x16x4f32 but row
    // (l/16)+4*q for v_mfma_f64_16x16x4f64 (measured on gfx950).
    V4 cpre[4][4];
    if (CPRE) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = col0 + wn * 64 + j * 16 + lr;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int rin = (sizeof(T) == 8) ? (lk + 4 * q) : (lk * 4 + q);
                    const int r = row0 + wm * 64 + i * 16 + rin;
                    cpre[i][j][q] = (r < g.M && c < g.N) ? g.C[(long)r * g.ldc + c] : (T)0;
                }
            }
    }

    const int nslab = (g.K + BK - 1) / BK;
    load_slab(0);
    store_slab(0);
    __syncthreads();

    for (int s = 0; s < nslab; ++s) {
        const int buf = s & 1;
        if (s + 1 < nslab) load_slab((s + 1) * BK);
        const T *as = As[buf];
        const T *bs = Bs[buf];
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = as[(kk + lk) * LDA_S + wm * 64 + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < 4; ++j) bf[j] = bs[(kk + lk) * LDB_S + wn * 64 + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = mfma16x16x4(af[i], bf[j], acc[i][j]);
        }
        if (s + 1 < nslab) store_slab(buf ^ 1);
        __syncthreads();
    }

    // epilogue: C = Cprefetched - acc
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = col0 + wn * 64 + j * 16 + lr;
            if (c >= g.N) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rin = (sizeof(T) == 8) ? (lk + 4 * q) : (lk * 4 + q);
                const int r = row0 + wm * 64 + i * 16 + rin;
                if (r < g.M) {
                    T *p = g.C + (long)r * g.ldc + c;
                    *p = (CPRE ? cpre[i][j][q] : *p) - acc[i][j][q];
                }
            }
        }
    }
}

// FP64, op(B) = B, K = KS * 16 (the LU's rank-128 update: KS = 8), two
// workgroups per CU.  C is never held beside the accumulators: A is negated on
// its way into LDS, so the MFMAs accumulate -A B, and each wave folds C in one
// block row at a time — the loads of block row i are issued after slab 2i's
// operand loads and added during slab 2i + 1, so they overlap the matrix-core
// work instead of preceding it (the memory counter retires loads in order: a
// wait for the slab operands never waits for C).  Operands come in as 16-byte
// loads; B goes to LDS as 16-byte stores.
// ASSIGN: C = A op(B) instead (C is not read; A enters LDS unnegated): the
// Cholesky TRSM L21 = A21 L11^-T, in place — a workgroup reads all K columns of
// its rows before it writes them, and with N <= 128 no other workgroup reads
// those rows.
template <int KS, bool VEC, int OCC, bool TRANS_B, bool ASSIGN = false>
__global__ __launch_bounds__(NTHREADS, OCC) void dgemm_fold_kernel(GemmArgs<double> g)
{
    static_assert(KS >= 8 && KS % 8 == 0, "eight half block rows of C folded over the slabs");
    if (g.info && *g.info != 0) return;
    int tm, tn;
    scirs_tile_of(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n, tm, tn);
    if (tm >= g.tiles_m || tn >= g.tiles_n) return;
    const int row0 = tm * BM, col0 = tn * BN;
    if (g.lower && col0 > row0 + BM - 1) return;   // SYRK: tiles on or below the diagonal

    __shared__ double As[2][BK * LDA_S];
    __shared__ double Bs[2][BK * LDB_S];

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int a_r = t >> 1, a_k = (t & 1) * 8;
    const int b_k = t >> 4, b_c = (t & 15) * 8;
    // Rows of A past M and columns of B past N only feed outputs that are never
    // stored, so their loads are clamped to a valid address instead of
    // predicated: no branch around a load, so the memory-counter waits stay exact.
    const double *pa0 = g.A + (long)min(row0 + a_r, g.M - 1) * g.lda + a_k;
    // op(B) = B^T: B is read like A, row col0 + a_r, eight consecutive k
    const double *pbt0 = g.B + (long)min(col0 + a_r, g.N - 1) * g.ldb + a_k;
    typedef int i4 __attribute__((ext_vector_type(4)));
    typedef double d8 __attribute__((ext_vector_type(8)));
    i4 bcol;   // first column of each of the thread's four B column pairs
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int c = col0 + b_c + 2 * i;
        bcol[i] = VEC ? (c + 1 < g.N ? c : g.N - 2) : c;
    }
    d8 ra, rb;   // vector types: the staging stays in registers

    auto load_slab = [&](int k0) {
        const double *pa = pa0 + k0;
        const double *pb = g.B + (long)(k0 + b_k) * g.ldb;
        const double *pbt = pbt0 + k0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double2 va, vb;
            if (VEC) {
                va = *reinterpret_cast<const double2 *>(pa + 2 * i);
                vb = TRANS_B ? *reinterpret_cast<const double2 *>(pbt + 2 * i)
                             : *reinterpret_cast<const double2 *>(pb + bcol[i]);
            } else {
                va = double2{pa[2 * i], pa[2 * i + 1]};
                vb = TRANS_B ? double2{pbt[2 * i], pbt[2 * i + 1]}
                             : double2{pb[min(bcol[i], g.N - 1)], pb[min(bcol[i] + 1, g.N - 1)]};
            }
            ra[2 * i] = va.x;
            ra[2 * i + 1] = va.y;
            rb[2 * i] = vb.x;
            rb[2 * i + 1] = vb.y;
        }
    };
    auto store_slab = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 8; ++i) As[buf][(a_k + i) * LDA_S + a_r] = ASSIGN ? ra[i] : -ra[i];
        if (TRANS_B) {
#pragma unroll
            for (int i = 0; i < 8; ++i) Bs[buf][(a_k + i) * LDB_S + a_r] = rb[i];
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                *reinterpret_cast<double2 *>(&Bs[buf][b_k * LDB_S + b_c + 2 * i]) =
                    double2{rb[2 * i], rb[2 * i + 1]};
        }
    };

    d4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = d4{0, 0, 0, 0};

    const int lr = lane & 15;
    const int lk = lane >> 4;
    // lane l, register q of a 16 x 16 f64 block: column l % 16, row l / 16 + 4 q
    // C arrives in eight half block rows h = 0..7 (block row h / 2, column
    // blocks 2 (h % 2) and 2 (h % 2) + 1), one per slab
    double cb[2][4];
    auto load_c = [&](int h) {
        const int i = h >> 1;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
            const int c = min(col0 + wn * 64 + (2 * (h & 1) + jj) * 16 + lr, g.N - 1);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = min(row0 + wm * 64 + i * 16 + lk + 4 * q, g.M - 1);
                cb[jj][q] = g.C[(long)r * g.ldc + c];
            }
        }
    };
    auto fold_c = [&](int h) {
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[h >> 1][2 * (h & 1) + jj][q] += cb[jj][q];
    };

    load_slab(0);
    store_slab(0);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int buf = s & 1;
        if (s + 1 < KS) load_slab((s + 1) * BK);
        // C half row h: loaded during slab h * KS / 8, folded in the next slab
        constexpr int STEP = KS / 8;
        if (!ASSIGN) {
            if (s >= STEP && s % STEP == 0) fold_c(s / STEP - 1);
            if (s % STEP == 0) load_c(s / STEP);
        }
        // keep the loads ahead of the matrix-core work and the LDS stores behind it
        __builtin_amdgcn_sched_barrier(0);
        const double *as = As[buf];
        const double *bs = Bs[buf];
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = as[(kk + lk) * LDA_S + wm * 64 + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < 4; ++j) bf[j] = bs[(kk + lk) * LDB_S + wn * 64 + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = mfma16x16x4(af[i], bf[j], acc[i][j]);
            // one k-step of fragments in registers at a time (the other wave on
            // the SIMD covers the LDS latency)
            __builtin_amdgcn_sched_barrier(0);
        }
        __builtin_amdgcn_sched_barrier(0);
        if (s + 1 < KS) store_slab(buf ^ 1);
        __syncthreads();
    }
    if (!ASSIGN) fold_c(7);
    // epilogue: C = C - A B (or A B), already in the accumulators
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = col0 + wn * 64 + j * 16 + lr;
            if (c >= g.N) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = row0 + wm * 64 + i * 16 + lk + 4 * q;
                if (r < g.M) g.C[(long)r * g.ldc + c] = acc[i][j][q];
            }
        }
    }
}

template <typename T>
int launch_gemm_sub(int M, int N, int K, const T *A, long lda, const T *B, long ldb, int trans_b, T *C,
                    long ldc, int lower, const int *info, hipStream_t stream)
{
    if (M <= 0 || N <= 0 || K <= 0) return SCIRS_OK;
    GemmArgs<T> g;
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
    g.tiles_m = (M + BM - 1) / BM;
    g.tiles_n = (N + BN - 1) / BN;
    g.lower = lower;
    g.info = info;
    const int nwg = (g.tiles_m * g.tiles_n + 7) / 8 * 8;
    // FP64: C prefetched at one workgroup per CU by default, which also leaves
    // room on each CU for the look-ahead panel running beside the update
    static const bool cpre = [] {
        const char *v = getenv("SCIRS_GPU_DGEMM_CPRE");
        return !(v && *v && atoi(v) == 0);
    }();
    static const int fold = [] {
        const char *v = getenv("SCIRS_GPU_DGEMM_FOLD");
        return v && *v ? atoi(v) : 1;
    }();
    if (sizeof(T) == 8 && fold && K == 8 * BK) {
        // the rank-128 update (LU) and SYRK (Cholesky): C folded into the
        // accumulators, two workgroups per CU
        const GemmArgs<double> &gd = reinterpret_cast<const GemmArgs<double> &>(g);
        // 16-byte operand loads: aligned rows and (B not transposed) column pairs
        // that are both in range or both out of it
        const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && lda % 2 == 0 && ldb % 2 == 0 &&
                         (trans_b || N % 2 == 0);
        if (trans_b) {
            if (vec)
                hipLaunchKernelGGL((dgemm_fold_kernel<8, true, 2, true>), dim3(nwg), dim3(NTHREADS), 0, stream, gd);
            else
                hipLaunchKernelGGL((dgemm_fold_kernel<8, false, 1, true>), dim3(nwg), dim3(NTHREADS), 0, stream, gd);
        } else if (vec) {
            hipLaunchKernelGGL((dgemm_fold_kernel<8, true, 2, false>), dim3(nwg), dim3(NTHREADS), 0, stream, gd);
        } else {
            hipLaunchKernelGGL((dgemm_fold_kernel<8, false, 1, false>), dim3(nwg), dim3(NTHREADS), 0, stream, gd);
        }
    } else if (trans_b)
        hipLaunchKernelGGL((gemm_sub_kernel<T, true, false>), dim3(nwg), dim3(NTHREADS), 0, stream, g);
    else if (cpre && sizeof(T) == 8)
        hipLaunchKernelGGL((gemm_sub_kernel<T, false, true>), dim3(nwg), dim3(NTHREADS), 0, stream, g);
    else
        hipLaunchKernelGGL((gemm_sub_kernel<T, false, false>), dim3(nwg), dim3(NTHREADS), 0, stream, g);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

} // namespace

int scirs_launch_dgemm_sub(int M, int N, int K, const double *A, long lda, const double *B, long ldb,
                           int trans_b, double *C, long ldc, int lower, const int *info,
                           hipStream_t stream)
{
    return launch_gemm_sub<double>(M, N, K, A, lda, B, ldb, trans_b, C, ldc, lower, info, stream);
}

int scirs_launch_dgemm_assign_t(int M, int N, int K, const double *A, long lda, const double *B, long ldb,
                                double *C, long ldc, const int *info, hipStream_t stream)
{
    if (M <= 0 || N <= 0) return SCIRS_OK;
    if (K != 8 * BK || N > BN) {
        scirs_set_error("dgemm_assign_t: K must be %d and N at most %d (got K=%d N=%d)", 8 * BK, BN, K, N);
        return SCIRS_ERR_ARG;
    }
    GemmArgs<double> g;
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
    g.tiles_m = (M + BM - 1) / BM;
    g.tiles_n = 1;
    g.lower = 0;
    g.info = info;
    const int nwg = (g.tiles_m + 7) / 8 * 8;
    const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && lda % 2 == 0 && ldb % 2 == 0;
    if (vec)
        hipLaunchKernelGGL((dgemm_fold_kernel<8, true, 2, true, true>), dim3(nwg), dim3(NTHREADS), 0, stream, g);
    else
        hipLaunchKernelGGL((dgemm_fold_kernel<8, false, 1, true, true>), dim3(nwg), dim3(NTHREADS), 0, stream, g);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_launch_sgemm_f32_sub(int M, int N, int K, const float *A, long lda, const float *B, long ldb,
                               int trans_b, float *C, long ldc, int lower, const int *info,
                               hipStream_t stream)
{
    return launch_gemm_sub<float>(M, N, K, A, lda, B, ldb, trans_b, C, ldc, lower, info, stream);
}

template <>
int scirs_trailing_update<double>(int M, int N, int K, const double *A, long lda, const double *B, long ldb,
                                  double *C, long ldc, const ScirsWork &w, hipStream_t stream)
{
    return scirs_launch_dgemm_sub(M, N, K, A, lda, B, ldb, 0, C, ldc, 0, w.info, stream);
}

template <>
int scirs_trailing_update<float>(int M, int N, int K, const float *A, long lda, const float *B, long ldb,
                                 float *C, long ldc, const ScirsWork &w, hipStream_t stream)
{
    const bool f16 = w.lowp_mode == SCIRS_LOWP_FP16 || w.lowp_mode == SCIRS_LOWP_FP16X3;
    if (f16 && w.colexp && N <= w.colexp_half) {
        // per-column scaling of U12; the side stream's update has its own region
        int *ce = w.colexp + (stream == w.side_stream ? w.colexp_half : 0);
        int rc = scirs_launch_colexp(K, N, B, ldb, ce, w.info, stream);
        if (rc) return rc;
        // |l_ij| <= 1 under partial pivoting: 2^14 lifts small multipliers
        return scirs_launch_sgemm_fp16_scaled(M, N, K, A, lda, B, ldb, C, ldc, ce, 14, w.lowp_mode, w.info, stream);
    }
    return scirs_launch_sgemm_sub(M, N, K, A, lda, B, ldb, 0, C, ldc, 0, w.lowp_mode, w.info, stream);
}

template <>
int scirs_syrk_update<double>(int N, int K, const double *A, long lda, double *C, long ldc, const ScirsWork &w,
                              hipStream_t stream)
{
    return scirs_launch_dgemm_sub(N, N, K, A, lda, A, lda, 1, C, ldc, 1, w.info, stream);
}

template <>
int scirs_syrk_update<float>(int N, int K, const float *A, long lda, float *C, long ldc, const ScirsWork &w,
                             hipStream_t stream)
{
    return scirs_launch_sgemm_sub(N, N, K, A, lda, A, lda, 1, C, ldc, 1, w.lowp_mode, w.info, stream);
}

template <>
int scirs_launch_gemm_sub_t<double>(int M, int N, int K, const double *A, long lda, const double *B, long ldb,
                                    double *C, long ldc, const ScirsWork &w, hipStream_t stream)
{
    return scirs_launch_dgemm_sub(M, N, K, A, lda, B, ldb, 1, C, ldc, 0, w.info, stream);
}

template <>
int scirs_launch_gemm_sub_t<float>(int M, int N, int K, const float *A, long lda, const float *B, long ldb, float *C,
                                   long ldc, const ScirsWork &w, hipStream_t stream)
{
    return scirs_launch_sgemm_sub(M, N, K, A, lda, B, ldb, 1, C, ldc, 0, w.lowp_mode, w.info, stream);
}
