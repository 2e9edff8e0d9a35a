// Bandwidth-bound helpers: FP64 residual R = B - A X, infinity norms,
// FP64 -> FP32 conversion of A fused with its row sums (||A||_inf), and the
// small vector updates of the refinement loops.  All row-major.
#include "common.hpp"
#include "kernels.hpp"
#include "prof.hpp"

namespace {

constexpr int NT = 256;

// R[i, j] = B[i, j] - sum_k A[i, k] X[k, j]   for j < nrhs (FP64 throughout).
// One wave per 4 rows; lanes stride over k so every A load is a coalesced
// 512-byte run; right-hand sides in chunks of 4.
__global__ __launch_bounds__(NT) void residual_kernel(int n, int nrhs, const double *A, long lda,
                                                      const double *X, long ldx, const double *B, long ldb,
                                                      double *R, long ldr)
{
    const int lane = threadIdx.x & 63;
    const int wave_global = (blockIdx.x * NT + threadIdx.x) >> 6;
    const int r0 = wave_global * 4;
    if (r0 >= n) return;
    const int nr = min(4, n - r0);
    for (int j0 = 0; j0 < nrhs; j0 += 4) {
        const int nj = min(4, nrhs - j0);
        double acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
        for (int k = lane; k < n; k += 64) {
            double xv[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) xv[b] = (b < nj) ? X[(long)k * ldx + j0 + b] : 0.0;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                if (a < nr) {
                    const double av = A[(long)(r0 + a) * lda + k];
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = fma(av, xv[b], acc[a][b]);
                }
            }
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = wave_sum(acc[a][b]);
        if (lane == 0) {
            for (int a = 0; a < nr; ++a)
                for (int b = 0; b < nj; ++b)
                    R[(long)(r0 + a) * ldr + j0 + b] = B[(long)(r0 + a) * ldb + j0 + b] - acc[a][b];
        }
    }
}

// One right-hand side with 16-byte aligned rows (even lda, aligned A and x):
// r = b - A x, one wave per 4 rows, each lane loading 2 consecutive doubles of
// 4 rows for 4 column chunks at once (32 loads in flight), x loaded once per
// chunk for the 4 rows.  The odd last column (n odd) is added by lane 0.
__global__ __launch_bounds__(NT) void residual1_kernel(int n, const double *A, long lda, const double *x,
                                                       const double *b, double *r)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    constexpr int U = 4;
    const int lane = threadIdx.x & 63;
    const int r0 = ((blockIdx.x * NT + threadIdx.x) >> 6) * 4;
    if (r0 >= n) return;
    const int nr = min(4, n - r0);
    const int ne = n & ~1;   // columns covered by 16-byte loads
    const double *rows[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) rows[a] = A + (long)(r0 + min(a, nr - 1)) * lda;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 2 * lane; k0 < ne; k0 += 128 * U) {
        d2 av[U][4], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = min(k0 + 128 * u, ne - 2);   // clamped: the surplus is masked below
            xv[u] = *reinterpret_cast<const d2 *>(x + k);
#pragma unroll
            for (int a = 0; a < 4; ++a) av[u][a] = *reinterpret_cast<const d2 *>(rows[a] + k);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool in = k0 + 128 * u < ne;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                acc[a] = fma(in ? av[u][a].x : 0.0, xv[u].x, acc[a]);
                acc[a] = fma(in ? av[u][a].y : 0.0, xv[u].y, acc[a]);
            }
        }
    }
    if (lane == 0 && ne < n)
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[a] = fma(rows[a][ne], x[ne], acc[a]);
#pragma unroll
    for (int a = 0; a < 4; ++a) acc[a] = wave_sum(acc[a]);
    if (lane == 0)
        for (int a = 0; a < nr; ++a) r[r0 + a] = b[r0 + a] - acc[a];
}

// out[j] = max_i |V[i, j]| for j < ncols (ncols <= 64); one workgroup per column.
template <typename T>
__global__ __launch_bounds__(NT) void colmax_kernel(int n, int ncols, const T *V, long ldv, double *out)
{
    const int j = blockIdx.x;
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) m = fmax(m, fabs((double)V[(long)i * ldv + j]));
    m = wave_max(m);
    __shared__ double s[NT / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = s[0];
        for (int w = 1; w < NT / 64; ++w) r = fmax(r, s[w]);
        out[j] = r;
    }
}

// A32 = fp32(A64) and rowsum[i] = sum_k |A64[i, k]| (one wave per row).
__global__ __launch_bounds__(NT) void convert_rowsum_kernel(int n, int m, const double *A, long lda, float *A32,
                                                            long lda32, double *rowsum)
{
    const int lane = threadIdx.x & 63;
    const int row = (blockIdx.x * NT + threadIdx.x) >> 6;
    if (row >= n) return;
    const double *a = A + (long)row * lda;
    float *o = A32 ? A32 + (long)row * lda32 : nullptr;
    double s = 0.0;
    for (int k = lane; k < m; k += 64) {
        const double v = a[k];
        if (o) o[k] = (float)v;
        s += fabs(v);
    }
    s = wave_sum(s);
    if (lane == 0 && rowsum) rowsum[row] = s;
}

// one workgroup: out = max_i v[i]
__global__ __launch_bounds__(NT) void max_kernel(int n, const double *v, double *out)
{
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) m = fmax(m, v[i]);
    m = wave_max(m);
    __shared__ double s[NT / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = s[0];
        for (int w = 1; w < NT / 64; ++w) r = fmax(r, s[w]);
        *out = r;
    }
}

template <typename S, typename D>
__global__ void convert_kernel(long count, int cols, const S *src, long lds, D *dst, long ldd)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    long i = idx / cols, j = idx % cols;
    dst[i * ldd + j] = (D)src[i * lds + j];
}

// X[i, j] += (double) D[i, j]
template <typename S>
__global__ void add_kernel(long count, int cols, const S *D, long ldd, double *X, long ldx)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    long i = idx / cols, j = idx % cols;
    X[i * ldx + j] += (double)D[i * ldd + j];
}

__global__ void fill_int_kernel(int *p, int count, int value)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

} // namespace

int scirs_residual(int n, int nrhs, const double *A, long lda, const double *X, long ldx, const double *B, long ldb,
                   double *R, long ldr, hipStream_t s)
{
    const int waves = (n + 3) / 4;
    const int blocks = (waves * 64 + NT - 1) / NT;
    scirs::ProfScope ps(SCIRS_PROF_RESIDUAL, 2.0 * n * n * nrhs, 8.0 * n * n + 24.0 * n * nrhs, s);
    const bool vec1 = nrhs == 1 && ldx == 1 && ldb == 1 && ldr == 1 && n >= 2 && lda % 2 == 0 &&
                      ((uintptr_t)A & 15) == 0 && ((uintptr_t)X & 15) == 0;
    if (vec1)
        hipLaunchKernelGGL(residual1_kernel, dim3(blocks), dim3(NT), 0, s, n, A, lda, X, B, R);
    else
        hipLaunchKernelGGL(residual_kernel, dim3(blocks), dim3(NT), 0, s, n, nrhs, A, lda, X, ldx, B, ldb, R, ldr);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_colmax_d(int n, int ncols, const double *V, long ldv, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(colmax_kernel<double>, dim3(ncols), dim3(NT), 0, s, n, ncols, V, ldv, out);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_colmax_s(int n, int ncols, const float *V, long ldv, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(colmax_kernel<float>, dim3(ncols), dim3(NT), 0, s, n, ncols, V, ldv, out);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_convert_rowsum(int n, int m, const double *A, long lda, float *A32, long lda32, double *rowsum,
                         hipStream_t s)
{
    const int blocks = (n * 64 + NT - 1) / NT;
    scirs::ProfScope ps(SCIRS_PROF_CONVERT, 0.0, (8.0 + (A32 ? 4.0 : 0.0)) * n * m, s);
    hipLaunchKernelGGL(convert_rowsum_kernel, dim3(blocks), dim3(NT), 0, s, n, m, A, lda, A32, lda32, rowsum);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_rowsum_max(int n, const double *rowsum, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(max_kernel, dim3(1), dim3(NT), 0, s, n, rowsum, out);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_d2s(int rows, int cols, const double *src, long lds, float *dst, long ldd, hipStream_t s)
{
    long count = (long)rows * cols;
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                       count, cols, src, lds, dst, ldd);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_s2d(int rows, int cols, const float *src, long lds, double *dst, long ldd, hipStream_t s)
{
    long count = (long)rows * cols;
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                       count, cols, src, lds, dst, ldd);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_d2d_copy(int rows, int cols, const double *src, long lds, double *dst, long ldd, hipStream_t s)
{
    long count = (long)rows * cols;
    hipLaunchKernelGGL((convert_kernel<double, double>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                       count, cols, src, lds, dst, ldd);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_d2d_copy(int rows, int cols, const float *src, long lds, float *dst, long ldd, hipStream_t s)
{
    long count = (long)rows * cols;
    hipLaunchKernelGGL((convert_kernel<float, float>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                       count, cols, src, lds, dst, ldd);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_add_s_to_d(int rows, int cols, const float *D, long ldd, double *X, long ldx, hipStream_t s)
{
    long count = (long)rows * cols;
    hipLaunchKernelGGL(add_kernel<float>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, count, cols, D,
                       ldd, X, ldx);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_add_d_to_d(int rows, int cols, const double *D, long ldd, double *X, long ldx, hipStream_t s)
{
    long count = (long)rows * cols;
    hipLaunchKernelGGL(add_kernel<double>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, count, cols, D,
                       ldd, X, ldx);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}

int scirs_fill_int(int *p, int count, int value, hipStream_t s)
{
    hipLaunchKernelGGL(fill_int_kernel, dim3((count + 255) / 256), dim3(256), 0, s, p, count, value);
    SCIRS_HIP_CHECK(hipGetLastError());
    return SCIRS_OK;
}
