// Batched array operations of scirs2-linalg/src/batch/mod.rs on the GPU:
//
//   scirs_gpu_{d,s}batch_matmul  <- batch_matmul  batch/mod.rs:62-107
//   scirs_gpu_{d,s}batch_matvec  <- batch_matvec  batch/mod.rs:139-172
//   scirs_gpu_{d,s}batch_add     <- batch_add     batch/mod.rs:220-291
//   scirs_gpu_{d,s}batch_sum     <- batch_sum     batch/mod.rs:326-336
//
// Every result is bit-identical to the reference's loops: each output element
// is accumulated by one thread, from +0, in the reference's index order, with
// the product rounded before the add (floating-point contraction is switched
// off in these kernels, as Rust never fuses a*b + c).  The batch is stored
// contiguously (batch x m x k, C order), so the batch of products A_p B is one
// (batch*m) x k by k x n product.
#include "api_internal.hpp"

using namespace scirs;

namespace {

// C (M x N) = A (M x K) B (K x N), row-major.  64 x 64 output tile per
// 256-thread workgroup, 4 x 4 outputs per thread, k staged through LDS in
// steps of 16; the k loop never runs past K, so no padding term is ever added.
constexpr int MT = 64, NT_ = 64, KT = 16;

template <typename T>
__global__ __launch_bounds__(256) void batch_matmul_kernel(const T *__restrict__ A, const T *__restrict__ B,
                                                           T *__restrict__ C, long M, int K, int N)
{
#pragma clang fp contract(off)
    __shared__ T As[MT][KT + 1];
    __shared__ T Bs[KT][NT_ + 1];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const long row0 = (long)blockIdx.x * MT;
    const int col0 = blockIdx.y * NT_;
    T acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = (T)0;
    for (int k0 = 0; k0 < K; k0 += KT) {
        const int kn = min(KT, K - k0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {   // A tile: 64 rows x 16 k
            const int r = u * 16 + tid / 16, q = tid % 16;
            const long gr = row0 + r;
            As[r][q] = (gr < M && q < kn) ? A[gr * K + k0 + q] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {   // B tile: 16 k x 64 columns
            const int q = u * 4 + tid / 64, cc = tid % 64;
            const int gc = col0 + cc;
            Bs[q][cc] = (q < kn && gc < N) ? B[(long)(k0 + q) * N + gc] : (T)0;
        }
        __syncthreads();
        for (int q = 0; q < kn; ++q) {
            T a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = As[ty * 4 + r][q];
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = Bs[q][tx + 16 * c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const T p = a[r] * b[c];
                    acc[r][c] = acc[r][c] + p;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const long gr = row0 + ty * 4 + r;
        if (gr >= M) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int gc = col0 + tx + 16 * c;
            if (gc < N) C[gr * N + gc] = acc[r][c];
        }
    }
}

// y (M) = A (M x N) x, one thread per row summing in column order; 16-column
// tiles of 256 rows staged through LDS so the loads are coalesced.
template <typename T>
__global__ __launch_bounds__(256) void batch_matvec_kernel(const T *__restrict__ A, const T *__restrict__ x,
                                                           T *__restrict__ y, long M, int N)
{
#pragma clang fp contract(off)
    __shared__ T As[256][17];
    __shared__ T xs[16];
    const int tid = threadIdx.x;
    const long row0 = (long)blockIdx.x * 256;
    T sum = (T)0;
    for (int j0 = 0; j0 < N; j0 += 16) {
        const int jn = min(16, N - j0);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int r = u * 16 + tid / 16, q = tid % 16;
            const long gr = row0 + r;
            As[r][q] = (gr < M && q < jn) ? A[gr * N + j0 + q] : (T)0;
        }
        if (tid < 16) xs[tid] = tid < jn ? x[j0 + tid] : (T)0;
        __syncthreads();
        for (int q = 0; q < jn; ++q) {
            const T p = As[tid][q] * xs[q];
            sum = sum + p;
        }
        __syncthreads();
    }
    if (row0 + tid < M) y[row0 + tid] = sum;
}

// out = a + v broadcast along rows (axis 0: v[i]) or columns (axis 1: v[j])
template <typename T>
__global__ __launch_bounds__(256) void batch_add_kernel(const T *__restrict__ a, const T *__restrict__ v,
                                                        T *__restrict__ out, long total, int m, int n, int axis)
{
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (long)gridDim.x * blockDim.x) {
        const int j = (int)(o % n);
        const int i = (int)((o / n) % m);
        out[o] = a[o] + (axis == 0 ? v[i] : v[j]);
    }
}

// out[o] = ((0 + a[0][o]) + a[1][o]) + ... in batch order
template <typename T>
__global__ __launch_bounds__(256) void batch_sum_kernel(const T *__restrict__ a, T *__restrict__ out, int batch,
                                                        long mn)
{
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < mn; o += (long)gridDim.x * blockDim.x) {
        T s = (T)0;
        for (int p = 0; p < batch; ++p) s = s + a[(long)p * mn + o];
        out[o] = s;
    }
}

unsigned grid_for(long total)
{
    const long g = (total + 255) / 256;
    return (unsigned)std::max(1L, std::min(g, 65536L));
}

template <typename T>
int launch_matmul(int batch, int m, int k, int n, const T *dA, const T *dB, T *dC, hipStream_t s)
{
    const long M = (long)batch * m;
    const long mt = (M + MT - 1) / MT, nt = (n + NT_ - 1) / NT_;
    if (mt >= (1L << 24) || nt > 65535) {   // 2^32 work-items along x; y kept under 2^16 blocks
        scirs_set_error("batch_matmul: batch*m = %ld or n = %d too large", M, n);
        return SCIRS_ERR_ARG;
    }
    dim3 grid((unsigned)mt, (unsigned)nt);
    batch_matmul_kernel<T><<<grid, 256, 0, s>>>(dA, dB, dC, M, k, n);
    SCIRS_HIP_CHECK(hipGetLastError());
    return 0;
}

// Host-memory entry: copy in, run, copy out on the library stream.
struct Scratch3 {
    void *a, *b, *c;
};

int scratch3(size_t abytes, size_t bbytes, size_t cbytes, Scratch3 *out)
{
    Context &c = ctx();
    int rc;
    if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, std::max<size_t>(abytes, 256)))) return rc;
    const size_t boff = (bbytes + 255) / 256 * 256;
    if ((rc = c.ensure(&c.scratch_b, &c.scratch_b_bytes, boff + std::max<size_t>(cbytes, 256)))) return rc;
    out->a = c.scratch_a;
    out->b = c.scratch_b;
    out->c = (char *)c.scratch_b + boff;
    return 0;
}

#define BATCH_ENTER()                                                   \
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);                  \
    Context &c = ctx();                                                  \
    {                                                                    \
        int _rc = c.init(0);                                             \
        if (_rc) return finish(info, _rc, t0);                           \
    }                                                                    \
    hipStream_t s = c.stream;

template <typename T>
int host_batch_matmul(int batch, int m, int k, int n, const T *a, const T *b, T *out, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    if (batch < 0 || m < 0 || k < 0 || n < 0 || (batch * (long)m * k > 0 && !a) || ((long)k * n > 0 && !b) ||
        (batch * (long)m * n > 0 && !out)) {
        scirs_set_error("batch_matmul: invalid arguments");
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    const size_t mn = (size_t)batch * m * n;
    if (mn == 0) return finish(info, SCIRS_OK, t0);
    BATCH_ENTER();
    Scratch3 w;
    int rc;
    const size_t ab = (size_t)batch * m * k * sizeof(T), bb = (size_t)k * n * sizeof(T), cb = mn * sizeof(T);
    if ((rc = scratch3(ab, bb, cb, &w))) return finish(info, rc, t0);
    if (ab && (rc = scirs::copy_h2d_flat(w.a, a, ab, s))) return finish(info, rc, t0);
    if (bb && (rc = scirs::copy_h2d_flat(w.b, b, bb, s))) return finish(info, rc, t0);
    if ((rc = launch_matmul<T>(batch, m, k, n, (const T *)w.a, (const T *)w.b, (T *)w.c, s)))
        return finish(info, rc, t0);
    if ((rc = scirs::copy_d2h_flat(out, w.c, cb, s))) return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    return finish(info, SCIRS_OK, t0);
}

template <typename T>
int host_batch_matvec(int batch, int m, int n, const T *a, const T *x, T *y, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    if (batch < 0 || m < 0 || n < 0 || (batch * (long)m * n > 0 && (!a || !x)) || (batch * (long)m > 0 && !y)) {
        scirs_set_error("batch_matvec: invalid arguments");
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    const long M = (long)batch * m;
    if (M == 0) return finish(info, SCIRS_OK, t0);
    if ((M + 255) / 256 >= (1L << 24)) {
        scirs_set_error("batch_matvec: batch*m = %ld too large", M);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    BATCH_ENTER();
    Scratch3 w;
    int rc;
    const size_t ab = (size_t)M * n * sizeof(T), xb = (size_t)n * sizeof(T), yb = (size_t)M * sizeof(T);
    if ((rc = scratch3(ab, xb, yb, &w))) return finish(info, rc, t0);
    if (ab && (rc = scirs::copy_h2d_flat(w.a, a, ab, s))) return finish(info, rc, t0);
    if (xb && (rc = scirs::copy_h2d_flat(w.b, x, xb, s))) return finish(info, rc, t0);
    batch_matvec_kernel<T><<<(unsigned)((M + 255) / 256), 256, 0, s>>>((const T *)w.a, (const T *)w.b, (T *)w.c, M,
                                                                       n);
    SCIRS_HIP_CHECK(hipGetLastError());
    if ((rc = scirs::copy_d2h_flat(y, w.c, yb, s))) return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    return finish(info, SCIRS_OK, t0);
}

template <typename T>
int host_batch_add(int batch, int m, int n, const T *a, const T *v, int axis, T *out, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    const long total = (long)batch * m * n;
    if (batch < 0 || m < 0 || n < 0 || (axis != 0 && axis != 1) || (total > 0 && (!a || !v || !out))) {
        scirs_set_error("batch_add: invalid arguments (axis must be 0 or 1)");
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    if (total == 0) return finish(info, SCIRS_OK, t0);
    BATCH_ENTER();
    Scratch3 w;
    int rc;
    const size_t ab = (size_t)total * sizeof(T), vb = (size_t)(axis == 0 ? m : n) * sizeof(T);
    if ((rc = scratch3(ab, vb, ab, &w))) return finish(info, rc, t0);
    if ((rc = scirs::copy_h2d_flat(w.a, a, ab, s))) return finish(info, rc, t0);
    if ((rc = scirs::copy_h2d_flat(w.b, v, vb, s))) return finish(info, rc, t0);
    batch_add_kernel<T><<<grid_for(total), 256, 0, s>>>((const T *)w.a, (const T *)w.b, (T *)w.c, total, m, n, axis);
    SCIRS_HIP_CHECK(hipGetLastError());
    if ((rc = scirs::copy_d2h_flat(out, w.c, ab, s))) return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    return finish(info, SCIRS_OK, t0);
}

template <typename T>
int host_batch_sum(int batch, int m, int n, const T *a, T *out, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    const long mn = (long)m * n;
    if (batch < 0 || m < 0 || n < 0 || (batch * mn > 0 && !a) || (mn > 0 && !out)) {
        scirs_set_error("batch_sum: invalid arguments");
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    if (mn == 0) return finish(info, SCIRS_OK, t0);
    BATCH_ENTER();
    Scratch3 w;
    int rc;
    const size_t ab = (size_t)batch * mn * sizeof(T), ob = (size_t)mn * sizeof(T);
    if ((rc = scratch3(ab, 0, ob, &w))) return finish(info, rc, t0);
    if (ab && (rc = scirs::copy_h2d_flat(w.a, a, ab, s))) return finish(info, rc, t0);
    batch_sum_kernel<T><<<grid_for(mn), 256, 0, s>>>((const T *)w.a, (T *)w.c, batch, mn);
    SCIRS_HIP_CHECK(hipGetLastError());
    if ((rc = scirs::copy_d2h_flat(out, w.c, ob, s))) return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    return finish(info, SCIRS_OK, t0);
}

} // namespace

extern "C" int scirs_gpu_dbatch_matmul(int batch, int m, int k, int n, const double *a, const double *b, double *c,
                                       scirs_solve_info *info)
{
    return host_batch_matmul<double>(batch, m, k, n, a, b, c, info);
}

extern "C" int scirs_gpu_sbatch_matmul(int batch, int m, int k, int n, const float *a, const float *b, float *c,
                                       scirs_solve_info *info)
{
    return host_batch_matmul<float>(batch, m, k, n, a, b, c, info);
}

extern "C" int scirs_gpu_dbatch_matmul_dev(int batch, int m, int k, int n, const double *dA, const double *dB,
                                           double *dC, void *stream)
{
    if (batch < 0 || m < 0 || k < 0 || n < 0) return SCIRS_ERR_ARG;
    if ((long)batch * m * n == 0) return SCIRS_OK;
    if (!dC || ((long)batch * m * k > 0 && (!dA || !dB))) return SCIRS_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    int rc = ctx().init(0);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;   // NULL: the default stream, ordered after the caller's work on it
    return launch_matmul<double>(batch, m, k, n, dA, dB, dC, s);
}

extern "C" int scirs_gpu_dbatch_matvec(int batch, int m, int n, const double *a, const double *x, double *y,
                                       scirs_solve_info *info)
{
    return host_batch_matvec<double>(batch, m, n, a, x, y, info);
}

extern "C" int scirs_gpu_sbatch_matvec(int batch, int m, int n, const float *a, const float *x, float *y,
                                       scirs_solve_info *info)
{
    return host_batch_matvec<float>(batch, m, n, a, x, y, info);
}

extern "C" int scirs_gpu_dbatch_add(int batch, int m, int n, const double *a, const double *v, int axis,
                                    double *out, scirs_solve_info *info)
{
    return host_batch_add<double>(batch, m, n, a, v, axis, out, info);
}

extern "C" int scirs_gpu_sbatch_add(int batch, int m, int n, const float *a, const float *v, int axis, float *out,
                                    scirs_solve_info *info)
{
    return host_batch_add<float>(batch, m, n, a, v, axis, out, info);
}

extern "C" int scirs_gpu_dbatch_sum(int batch, int m, int n, const double *a, double *out, scirs_solve_info *info)
{
    return host_batch_sum<double>(batch, m, n, a, out, info);
}

extern "C" int scirs_gpu_sbatch_sum(int batch, int m, int n, const float *a, float *out, scirs_solve_info *info)
{
    return host_batch_sum<float>(batch, m, n, a, out, info);
}
