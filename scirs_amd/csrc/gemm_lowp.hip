// Low-precision Schur-complement update of the FP32 factorisation:
//     C[M x N] (FP32) -= A[M x K] * op(B)   with FP32 storage everywhere.
// The operands are rounded to 16-bit floats while they are staged into LDS
// and multiplied on the gfx950 v_mfma_f32_16x16x32_{bf16,f16} matrix cores
// with FP32 accumulation; C stays FP32.
//   SCIRS_LOWP_BF16    one BF16 product (u = 2^-8 per operand)
//   SCIRS_LOWP_FP16    one FP16 product; A is pre-scaled by 2^a_shift and
//                      B column c by 2^-colexp[c] (powers of two, undone exactly
//                      in the epilogue) so that neither operand leaves FP16 range
//   SCIRS_LOWP_BF16X3  x = hi + lo with hi = bf16(x), lo = bf16(x - hi);
//                      C -= Ahi Bhi + Ahi Blo + Alo Bhi  (about 16 significant
//                      bits per operand, three matrix-core products)
//   SCIRS_LOWP_FP16X3  the FP16 scaling above, then x = hi + lo with hi = f16(x),
//                      lo = f16(x - hi); C -= Ahi Bhi + Ahi Blo + Alo Bhi (about
//                      22 significant bits per operand on FP16 matrix cores)
//   SCIRS_LOWP_BF16X6  x = hi + mid + lo (three BF16 terms, ~24 bits: the FP32
//                      significand); the six products whose order is above
//                      2^-24 (hh, hm, mh, hl, lh, mm), so the update is as
//                      accurate as an FP32 one while running on BF16 cores
// Tiling: 128 x 128 per 256-thread workgroup, 2 x 2 waves of 64 x 64, K slabs
// of 32 (one MFMA k-step) double-buffered in LDS (single-buffered for the
// three-term split, whose staging is 1.5x larger).
#include "common.hpp"
#include "kernels.hpp"
#include "tile_map.hpp"

namespace {

constexpr int BM = 128, BN = 128, BK = 32;
constexpr int PITCH = BK + 8;   // 16-bit elements per LDS row (80 B: keeps 16 B alignment)
constexpr int NTHREADS = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

__device__ __forceinline__ u16 f32_to_bf16_rne(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(u16 h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ u16 f32_to_f16(float x)
{
    _Float16 h = (_Float16)x;
    return *reinterpret_cast<u16 *>(&h);
}

template <int MODE>
struct LowpArgs {
    int M, N, K;
    const float *A; long lda;
    const float *B; long ldb;
    float *C; long ldc;
    int tiles_m, tiles_n;
    int lower;
    const int *info;
    const int *colexp;   // FP16: per-column exponent of B (may be null)
    int a_shift;         // FP16: A is multiplied by 2^a_shift before rounding
};

template <int MODE, bool TRANS_B, bool VEC>
__global__ __launch_bounds__(NTHREADS, 2) void lowp_gemm_kernel(LowpArgs<MODE> g)
{
    if (g.info && *g.info != 0) return;
    int tm, tn;
    scirs_tile_of(blockIdx.x, gridDim.x, g.tiles_m, g.tiles_n, tm, tn);
    if (tm >= g.tiles_m || tn >= g.tiles_n) return;
    const int row0 = tm * BM, col0 = tn * BN;
    if (g.lower && col0 > row0 + BM - 1) return;

    constexpr bool F16 = MODE == SCIRS_LOWP_FP16 || MODE == SCIRS_LOWP_FP16X3;   // FP16 cores, scaled operands
    constexpr int NSPLIT =
        (MODE == SCIRS_LOWP_BF16X6) ? 3 : (MODE == SCIRS_LOWP_BF16X3 || MODE == SCIRS_LOWP_FP16X3) ? 2 : 1;
    constexpr int NBUF = NSPLIT == 3 ? 1 : 2;
    __shared__ __align__(16) u16 As[NBUF][NSPLIT][BM * PITCH];
    __shared__ __align__(16) u16 Bs[NBUF][NSPLIT][BN * PITCH];

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    // A slab: 128 rows x 32 k; thread t loads row t/2, k (t%2)*16 .. +16
    const int a_r = t >> 1, a_k = (t & 1) * 16;
    // B slab (row-major K x N): thread t loads k = t%32, cols (t/32)*16 .. +16
    const int b_k = t & 31, b_c = (t >> 5) * 16;
    float ra[16], rb[16];
    const float ascale = F16 ? __builtin_ldexpf(1.0f, g.a_shift) : 1.0f;

    // 16 consecutive floats of one row: four 16-byte loads when the row segment
    // is in range and aligned (VEC: every leading dimension a multiple of 4)
    auto load16 = [&](float (&dst)[16], const float *src, bool row_ok, int first, int limit) {
        if (VEC && row_ok && first + 15 < limit) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const f4 x = *reinterpret_cast<const f4 *>(src + 4 * v);
                dst[4 * v] = x[0]; dst[4 * v + 1] = x[1]; dst[4 * v + 2] = x[2]; dst[4 * v + 3] = x[3];
            }
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) dst[i] = (row_ok && first + i < limit) ? src[i] : 0.0f;
        }
    };
    auto load_slab = [&](int k0) {
        const int gr = row0 + a_r;
        load16(ra, g.A + (long)gr * g.lda + k0 + a_k, gr < g.M, k0 + a_k, g.K);
        if (!TRANS_B) {
            const int gk = k0 + b_k;
            load16(rb, g.B + (long)gk * g.ldb + col0 + b_c, gk < g.K, col0 + b_c, g.N);
        } else {
            const int gc = col0 + a_r;
            load16(rb, g.B + (long)gc * g.ldb + k0 + a_k, gc < g.N, k0 + a_k, g.K);
        }
    };
    auto cvt = [&](float x, u16 &hi, u16 &lo, u16 &mid) {
        if (F16) {
            hi = f32_to_f16(x);
            if (MODE == SCIRS_LOWP_FP16X3) {
                _Float16 h = *reinterpret_cast<const _Float16 *>(&hi);
                lo = f32_to_f16(x - (float)h);
            }
        } else {
            hi = f32_to_bf16_rne(x);
            if (MODE == SCIRS_LOWP_BF16X3) lo = f32_to_bf16_rne(x - bf16_to_f32(hi));
            if (MODE == SCIRS_LOWP_BF16X6) {
                const float r = x - bf16_to_f32(hi);
                mid = f32_to_bf16_rne(r);
                lo = f32_to_bf16_rne(r - bf16_to_f32(mid));
            }
        }
    };
    // split index 0 = hi, NSPLIT-1 = lo, 1 = mid (three-term split only)
    auto put = [&](u16 (*dst)[BM * PITCH], int idx, const u16 &hi, const u16 &lo, const u16 &mid) {
        dst[0][idx] = hi;
        if (NSPLIT >= 2) dst[NSPLIT - 1][idx] = lo;
        if (NSPLIT == 3) dst[1][idx] = mid;
    };
    auto store_slab = [&](int buf) {
        u16 hi[16], lo[16], mid[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) cvt(ra[i] * ascale, hi[i], lo[i], mid[i]);
#pragma unroll
        for (int i = 0; i < 16; ++i) put(As[buf], a_r * PITCH + a_k + i, hi[i], lo[i], mid[i]);
        if (!TRANS_B) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                float x = rb[i];
                if (F16 && g.colexp) {
                    const int gc = col0 + b_c + i;
                    x = __builtin_ldexpf(x, gc < g.N ? -g.colexp[gc] : 0);
                }
                cvt(x, hi[i], lo[i], mid[i]);
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) put(Bs[buf], (b_c + i) * PITCH + b_k, hi[i], lo[i], mid[i]);
        } else {
            float sc = 1.0f;
            if (F16 && g.colexp) {
                const int gc = col0 + a_r;
                sc = __builtin_ldexpf(1.0f, gc < g.N ? -g.colexp[gc] : 0);
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) cvt(rb[i] * sc, hi[i], lo[i], mid[i]);
#pragma unroll
            for (int i = 0; i < 16; ++i) put(Bs[buf], a_r * PITCH + a_k + i, hi[i], lo[i], mid[i]);
        }
    };

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    // The matrix-core products are formed as (op(B)^T A^T) so that lane l of a
    // 16 x 16 block holds row l%16 and the FOUR CONSECUTIVE columns 4*(l/16)+q:
    // C moves as 16-byte vectors.  C is read once; its loads are issued now so
    // their latency overlaps the K loop.
    const int lc = (lane >> 4) * 4;
    f4 cpre[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = row0 + wm * 64 + i * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = col0 + wn * 64 + j * 16 + lc;
            const float *src = g.C + (long)r * g.ldc + c;
            if (VEC && r < g.M && c + 3 < g.N) {
                cpre[i][j] = *reinterpret_cast<const f4 *>(src);
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) cpre[i][j][q] = (r < g.M && c + q < g.N) ? src[q] : 0.0f;
            }
        }
    }

    const int nslab = (g.K + BK - 1) / BK;
    load_slab(0);
    store_slab(0);
    __syncthreads();

    const int lr = lane & 15;        // row of A / column of B inside a 16-block
    const int lk = (lane >> 4) * 8;  // first of the lane's 8 k values

    for (int s = 0; s < nslab; ++s) {
        const int buf = NBUF == 2 ? (s & 1) : 0;
        if (s + 1 < nslab) load_slab((s + 1) * BK);
        typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
        u16x8 ah[4], bh[4], al[4], bl[4], am[4], bm[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int ra_ = (wm * 64 + i * 16 + lr) * PITCH + lk, rb_ = (wn * 64 + i * 16 + lr) * PITCH + lk;
            ah[i] = *reinterpret_cast<const u16x8 *>(&As[buf][0][ra_]);
            bh[i] = *reinterpret_cast<const u16x8 *>(&Bs[buf][0][rb_]);
            if (NSPLIT >= 2) {
                al[i] = *reinterpret_cast<const u16x8 *>(&As[buf][NSPLIT - 1][ra_]);
                bl[i] = *reinterpret_cast<const u16x8 *>(&Bs[buf][NSPLIT - 1][rb_]);
            }
            if (NSPLIT == 3) {
                am[i] = *reinterpret_cast<const u16x8 *>(&As[buf][1][ra_]);
                bm[i] = *reinterpret_cast<const u16x8 *>(&Bs[buf][1][rb_]);
            }
        }
        auto mma = [&](const u16x8 &x, const u16x8 &y, f4 &c) {
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, x), __builtin_bit_cast(bf16x8, y),
                                                        c, 0, 0, 0);
        };
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                // operands swapped: op(B) fragment as the left factor, A as the right
                if (F16) {
                    auto hma = [&](const u16x8 &x, const u16x8 &y, f4 &c) {
                        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, x),
                                                                   __builtin_bit_cast(f16x8, y), c, 0, 0, 0);
                    };
                    if (MODE == SCIRS_LOWP_FP16X3) {
                        hma(bl[j], ah[i], acc[i][j]);
                        hma(bh[j], al[i], acc[i][j]);
                    }
                    hma(bh[j], ah[i], acc[i][j]);
                } else if (MODE == SCIRS_LOWP_BF16X6) {
                    // smallest terms first
                    mma(bm[j], am[i], acc[i][j]);
                    mma(bl[j], ah[i], acc[i][j]);
                    mma(bh[j], al[i], acc[i][j]);
                    mma(bm[j], ah[i], acc[i][j]);
                    mma(bh[j], am[i], acc[i][j]);
                    mma(bh[j], ah[i], acc[i][j]);
                } else {
                    mma(bh[j], ah[i], acc[i][j]);
                    if (MODE == SCIRS_LOWP_BF16X3) {
                        mma(bl[j], ah[i], acc[i][j]);
                        mma(bh[j], al[i], acc[i][j]);
                    }
                }
            }
        if (s + 1 < nslab) {
            if (NBUF == 1) __syncthreads();   // every wave is done with the single buffer
            store_slab(NBUF == 2 ? (buf ^ 1) : 0);
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = row0 + wm * 64 + i * 16 + (lane & 15);
        if (r >= g.M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = col0 + wn * 64 + j * 16 + lc;
            f4 out;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float sc = 1.0f;
                if (F16)
                    sc = __builtin_ldexpf(1.0f, ((g.colexp && c + q < g.N) ? g.colexp[c + q] : 0) - g.a_shift);
                out[q] = cpre[i][j][q] - acc[i][j][q] * sc;
            }
            float *dst = g.C + (long)r * g.ldc + c;
            if (VEC && c + 3 < g.N) {
                *reinterpret_cast<f4 *>(dst) = out;
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (c + q < g.N) dst[q] = out[q];
            }
        }
    }
}

template <int MODE>
int launch_lowp(int M, int N, int K, const float *A, long lda, const float *B, long ldb, int trans_b,
                float *C, long ldc, int lower, const int *info, const int *colexp, int a_shift,
                hipStream_t stream)
{
    LowpArgs<MODE> g;
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
    g.tiles_m = (M + BM - 1) / BM;
    g.tiles_n = (N + BN - 1) / BN;
    g.lower = lower; g.info = info; g.colexp = colexp; g.a_shift = a_shift;
    const int nwg = (g.tiles_m * g.tiles_n + 7) / 8 * 8;
    auto al16 = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
    const bool vec = lda % 4 == 0 && ldb % 4 == 0 && ldc % 4 == 0 && al16(A) && al16(B) && al16(C);
    if (trans_b) {
        if (vec) hipLaunchKernelGGL((lowp_gemm_kernel<MODE, true, true>), dim3(nwg), dim3(NTHREADS), 0, stream, g);
        else hipLaunchKernelGGL((lowp_gemm_kernel<MODE, true, false>), dim3(nwg), dim3(NTHREADS), 0, stream, g);
    } else {
        if (vec) hipLaunchKernelGGL((lowp_gemm_kernel<MODE, false, true>), dim3(nwg), dim3(NTHREADS), 0, stream, g);
        else hipLaunchKernelGGL((lowp_gemm_kernel<MODE, false, false>), dim3(nwg), dim3(NTHREADS), 0, stream, g);
    }
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

} // namespace

int scirs_launch_sgemm_f32_sub(int M, int N, int K, const float *A, long lda, const float *B, long ldb,
                               int trans_b, float *C, long ldc, int lower, const int *info,
                               hipStream_t stream);

int scirs_launch_sgemm_sub(int M, int N, int K, const float *A, long lda, const float *B, long ldb,
                           int trans_b, float *C, long ldc, int lower, int mode, const int *info,
                           hipStream_t stream)
{
    if (M <= 0 || N <= 0 || K <= 0) return SCIRS_OK;
    switch (mode) {
    case SCIRS_LOWP_FP32:
        return scirs_launch_sgemm_f32_sub(M, N, K, A, lda, B, ldb, trans_b, C, ldc, lower, info, stream);
    case SCIRS_LOWP_BF16:
        return launch_lowp<SCIRS_LOWP_BF16>(M, N, K, A, lda, B, ldb, trans_b, C, ldc, lower, info, nullptr, 0, stream);
    case SCIRS_LOWP_BF16X3:
        return launch_lowp<SCIRS_LOWP_BF16X3>(M, N, K, A, lda, B, ldb, trans_b, C, ldc, lower, info, nullptr, 0, stream);
    case SCIRS_LOWP_BF16X6:
        return launch_lowp<SCIRS_LOWP_BF16X6>(M, N, K, A, lda, B, ldb, trans_b, C, ldc, lower, info, nullptr, 0, stream);
    case SCIRS_LOWP_FP16:
        // |l_ij| <= 1 under partial pivoting: 2^14 keeps A in range and lifts small multipliers
        return launch_lowp<SCIRS_LOWP_FP16>(M, N, K, A, lda, B, ldb, trans_b, C, ldc, lower, info, nullptr, 14, stream);
    case SCIRS_LOWP_FP16X3:
        return launch_lowp<SCIRS_LOWP_FP16X3>(M, N, K, A, lda, B, ldb, trans_b, C, ldc, lower, info, nullptr, 14,
                                              stream);
    default:
        scirs_set_error("unknown low-precision mode %d", mode);
        return SCIRS_ERR_ARG;
    }
}

namespace {
// Per-column power-of-two exponent of the FP16 update's right operand: column c
// of B (K x N, row-major) is multiplied by 2^-colexp[c] before rounding to FP16,
// which puts its largest entry in [2^14, 2^15) (FP16 tops out below 2^16), and
// the product is multiplied back by 2^colexp[c] in FP32: exact scalings.
__global__ void colexp_kernel(int K, int N, const float *B, long ldb, int *colexp, const int *info)
{
    if (info && *info != 0) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float m = 0.0f;
    for (int k = 0; k < K; ++k) m = fmaxf(m, fabsf(B[(long)k * ldb + c]));
    colexp[c] = (m > 0.0f && m < INFINITY) ? ilogbf(m) - 14 : 0;
}
} // namespace

int scirs_launch_colexp(int K, int N, const float *B, long ldb, int *colexp, const int *info, hipStream_t stream)
{
    if (N <= 0) return SCIRS_OK;
    hipLaunchKernelGGL(colexp_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, K, N, B, ldb, colexp, info);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_launch_sgemm_fp16_scaled(int M, int N, int K, const float *A, long lda, const float *B, long ldb,
                                   float *C, long ldc, const int *colexp, int a_shift, int mode, const int *info,
                                   hipStream_t stream)
{
    if (M <= 0 || N <= 0 || K <= 0) return SCIRS_OK;
    if (mode == SCIRS_LOWP_FP16X3)
        return launch_lowp<SCIRS_LOWP_FP16X3>(M, N, K, A, lda, B, ldb, 0, C, ldc, 0, info, colexp, a_shift, stream);
    return launch_lowp<SCIRS_LOWP_FP16>(M, N, K, A, lda, B, ldb, 0, C, ldc, 0, info, colexp, a_shift, stream);
}
