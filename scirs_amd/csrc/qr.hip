// Householder QR and the least-squares solve built on it.
//
//   scirs_gpu_dqr     <- qr_factor   scirs2-linalg/src/lapack.rs:171-253
//                        (qr          scirs2-linalg/src/decomposition.rs:208-260)
//   scirs_gpu_dlstsq  <- lstsq       scirs2-linalg/src/solve.rs:234-290 (rows >= cols)
//
// The factorisation follows the reference's reflector sequence: for column k
// the reflector v = x - alpha e1 (alpha = -||x|| when x0 >= 0, else +||x||) is
// normalised to unit length and H = I - 2 v v^T is applied to the trailing
// columns of R and to the rows k.. of the accumulated Q^T; a column whose
// ||x|| or ||v|| is not above machine epsilon is left as it is.  Per column
// three launches: one workgroup forms v (two block reductions in a fixed
// order); then the trailing block, cut into 32-column strips x 256-row chunks,
// forms per-chunk partials of v^T c and subtracts 2 v (v^T c) after summing
// the partials in chunk order.  Row-major storage makes each 32-column strip
// of a row a 256-byte coalesced segment.  This is the unblocked (level-2)
// form: bound by passes over the trailing block per column; the blocked
// (compact-WY, MFMA) form is the next step for large matrices.
#include "api_internal.hpp"
#include <cfloat>
#include <vector>

using namespace scirs;

namespace {

constexpr int HT = 1024;   // threads of the reflector kernel
constexpr int AC = 32;     // columns per apply workgroup
constexpr int AG = 8;      // row groups per apply workgroup
constexpr int AT = AC * AG;

// Sum over the workgroup, the same order on every call (waves, then wave sums).
__device__ double block_sum(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    return s;
}

// v[0 .. rows-k) <- the unit reflector of column k of R; flag <- 1, or 0 when
// the reference skips the column (lapack.rs:205-215).  The column is read
// once: its first HC*HT entries stay in registers for the three passes.
constexpr int HC = 16;

__global__ __launch_bounds__(HT) void house_kernel(const double *__restrict__ R, long ldr, int rows, int k,
                                                   double *__restrict__ v, int *__restrict__ flag)
{
    __shared__ double red[HT / 64];
    const int len = rows - k;
    const double *col = R + (long)k * ldr + k;
    double xc[HC];
    double ss = 0.0;
#pragma unroll
    for (int u = 0; u < HC; ++u) {
        const int i = threadIdx.x + u * HT;
        xc[u] = i < len ? col[(long)i * ldr] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < HC; ++u) ss += xc[u] * xc[u];
    for (int i = threadIdx.x + HC * HT; i < len; i += HT) {
        const double x = col[(long)i * ldr];
        ss += x * x;
    }
    const double xn = sqrt(block_sum(ss, red));
    if (!(xn > DBL_EPSILON)) {
        if (threadIdx.x == 0) *flag = 0;
        return;
    }
    const double x0 = col[0];
    const double alpha = x0 >= 0.0 ? -xn : xn;
    const double v0 = x0 - alpha;
    if (threadIdx.x == 0) xc[0] = v0;
    double vs = 0.0;
#pragma unroll
    for (int u = 0; u < HC; ++u) vs += xc[u] * xc[u];
    for (int i = threadIdx.x + HC * HT; i < len; i += HT) {
        const double x = col[(long)i * ldr];
        vs += x * x;
    }
    const double vn = sqrt(block_sum(vs, red));
    if (!(vn > DBL_EPSILON)) {
        if (threadIdx.x == 0) *flag = 0;
        return;
    }
#pragma unroll
    for (int u = 0; u < HC; ++u) {
        const int i = threadIdx.x + u * HT;
        if (i < len) v[i] = xc[u] / vn;
    }
    for (int i = threadIdx.x + HC * HT; i < len; i += HT) v[i] = col[(long)i * ldr] / vn;
    if (threadIdx.x == 0) *flag = 1;
}

// Apply H = I - 2 v v^T to rows k.. of two matrices: R's columns [rc0, rc1)
// (column strips [0, nbr)) and all qcols columns of Qt (the remaining strips).
// The trailing rows are split into chunks of RC rows (blockIdx.y), so the
// update fills the chip at any shape: pass 1 writes each chunk's partial
// v^T c per column, pass 2 sums the partials in chunk order (the same order
// in every workgroup) and subtracts 2 v (v^T c) from its chunk.
constexpr int RC = 256;

struct StripCol {
    double *M;
    long ld;
    int c, cend, g;   // column in M, its end, column index in the partials
};

__device__ StripCol strip_col(double *R, long ldr, int rc0, int rc1, int nbr, double *Qt, long ldq, int qcols)
{
    const int cl = threadIdx.x % AC;
    int b = blockIdx.x;
    StripCol s;
    s.g = b * AC + cl;
    if (b < nbr) {
        s.M = R; s.ld = ldr; s.c = rc0 + b * AC + cl; s.cend = rc1;
    } else {
        b -= nbr;
        s.M = Qt; s.ld = ldq; s.c = b * AC + cl; s.cend = qcols;
    }
    return s;
}

__global__ __launch_bounds__(AT) void house_dots_kernel(double *__restrict__ R, long ldr, int rc0, int rc1, int nbr,
                                                        double *__restrict__ Qt, long ldq, int qcols, int rows, int k,
                                                        const double *__restrict__ v, const int *__restrict__ flag,
                                                        double *__restrict__ part, long ldp)
{
    if (*flag == 0) return;
    __shared__ double red[AG][AC];
    const StripCol sc = strip_col(R, ldr, rc0, rc1, nbr, Qt, ldq, qcols);
    const int cl = threadIdx.x % AC, rg = threadIdx.x / AC;
    const int len = rows - k, i0 = blockIdx.y * RC, i1 = min(len, i0 + RC);
    double acc = 0.0;
    if (sc.c < sc.cend) {
        const double *base = sc.M + (long)k * sc.ld + sc.c;
        for (int i = i0 + rg; i < i1; i += AG) acc += v[i] * base[(long)i * sc.ld];
    }
    red[rg][cl] = acc;
    __syncthreads();
    if (rg == 0) {
        double d = 0.0;
#pragma unroll
        for (int g = 0; g < AG; ++g) d += red[g][cl];
        part[(long)blockIdx.y * ldp + sc.g] = d;
    }
}

__global__ __launch_bounds__(AT) void house_update_kernel(double *__restrict__ R, long ldr, int rc0, int rc1,
                                                          int nbr, double *__restrict__ Qt, long ldq, int qcols,
                                                          int rows, int k, const double *__restrict__ v,
                                                          const int *__restrict__ flag,
                                                          const double *__restrict__ part, long ldp, int nrc)
{
    if (*flag == 0) return;
    const StripCol sc = strip_col(R, ldr, rc0, rc1, nbr, Qt, ldq, qcols);
    if (sc.c >= sc.cend) return;
    const int rg = threadIdx.x / AC;
    double dot = 0.0;
    for (int q = 0; q < nrc; ++q) dot += part[(long)q * ldp + sc.g];
    const int len = rows - k, i0 = blockIdx.y * RC, i1 = min(len, i0 + RC);
    double *base = sc.M + (long)k * sc.ld + sc.c;
    for (int i = i0 + rg; i < i1; i += AG) base[(long)i * sc.ld] -= 2.0 * v[i] * dot;
}

__global__ void eye_kernel(double *__restrict__ Q, long ldq, int n)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * n) return;
    const int i = (int)(idx / n), j = (int)(idx % n);
    Q[(long)i * ldq + j] = i == j ? 1.0 : 0.0;
}

// dst (cols x rows) <- src^T (src rows x cols), 32 x 32 tiles through LDS.
__global__ __launch_bounds__(256) void transpose_kernel(const double *__restrict__ src, long lds, int rows, int cols,
                                                        double *__restrict__ dst, long ldd)
{
    __shared__ double t[32][33];
    const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    for (int y = ty; y < 32; y += 8)
        if (r0 + y < rows && c0 + tx < cols) t[y][tx] = src[(long)(r0 + y) * lds + c0 + tx];
    __syncthreads();
    for (int y = ty; y < 32; y += 8)
        if (c0 + y < cols && r0 + tx < rows) dst[(long)(c0 + y) * ldd + r0 + tx] = t[tx][y];
}

// sq[i] <- (b_i - A_i x)^2, one wave per row
__global__ __launch_bounds__(256) void resid_sq_kernel(const double *__restrict__ A, long lda, int rows, int cols,
                                                       const double *__restrict__ b, const double *__restrict__ x,
                                                       double *__restrict__ sq)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
    if (row >= rows) return;
    double acc = 0.0;
    for (int j = l; j < cols; j += 64) acc += A[(long)row * lda + j] * x[j];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (l == 0) {
        const double d = b[row] - acc;
        sq[row] = d * d;
    }
}

// out[i] = M[i * stride], i < count (a column or the diagonal of a row-major matrix)
__global__ __launch_bounds__(256) void strided_gather_kernel(const double *__restrict__ M, long stride, int count,
                                                             double *__restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) out[i] = M[(long)i * stride];
}

// M[i * stride] = in[i], i < count
__global__ __launch_bounds__(256) void strided_scatter_kernel(const double *__restrict__ in, int count,
                                                              double *__restrict__ M, long stride)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) M[(long)i * stride] = in[i];
}

__global__ __launch_bounds__(HT) void sum_kernel(const double *__restrict__ sq, int n, double *__restrict__ out)
{
    __shared__ double red[HT / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += HT) s += sq[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) *out = s;
}

// Householder sweep over the first min(rows, cols) columns of R (rows x rcols,
// ld ldr); Qt (rows x rows, ld ldq) accumulates Q^T when not null.  part
// holds partial_elems(rows, rcols, Qt != null) doubles.
long partial_elems(int rows, int rcols, bool with_q)
{
    const long strips = (rcols + AC - 1) / AC + (with_q ? (rows + AC - 1) / AC : 0);
    return ((rows + RC - 1) / RC) * strips * AC;
}

int householder_sweep(int rows, int cols, int rcols, double *R, long ldr, double *Qt, long ldq, double *v, int *flag,
                      double *part, hipStream_t s)
{
    const int kmax = std::min(rows, cols);
    for (int k = 0; k < kmax; ++k) {
        house_kernel<<<1, HT, 0, s>>>(R, ldr, rows, k, v, flag);
        const int nbr = (rcols - k + AC - 1) / AC;
        const int nbq = Qt ? (rows + AC - 1) / AC : 0;
        const int nrc = (rows - k + RC - 1) / RC;
        const long ldp = (long)(nbr + nbq) * AC;
        dim3 grid(nbr + nbq, nrc);
        house_dots_kernel<<<grid, AT, 0, s>>>(R, ldr, k, rcols, nbr, Qt, ldq, rows, rows, k, v, flag, part, ldp);
        house_update_kernel<<<grid, AT, 0, s>>>(R, ldr, k, rcols, nbr, Qt, ldq, rows, rows, k, v, flag, part, ldp,
                                                nrc);
    }
    SCIRS_HIP_CHECK(hipGetLastError());
    return 0;
}

long round_up(long x, long m) { return (x + m - 1) / m * m; }

} // namespace

extern "C" int scirs_gpu_dqr(int rows, int cols, const double *a, int64_t lda, double *q, int64_t ldq, double *r,
                             int64_t ldr, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    if (rows <= 0 || cols <= 0 || rows < cols || !a || !q || !r || lda < cols || ldq < rows || ldr < cols) {
        scirs_set_error("dqr: invalid arguments (rows=%d cols=%d; rows >= cols >= 1)", rows, cols);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    Context &c = ctx();
    int rc = c.init(0);
    if (rc) return finish(info, rc, t0);
    hipStream_t s = c.stream;
    const long pr = round_up(cols, 64), pq = round_up(rows, 64);
    const size_t rbytes = (size_t)rows * pr * sizeof(double), qbytes = (size_t)rows * pq * sizeof(double);
    if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, rbytes))) return finish(info, rc, t0);
    const size_t npart = (size_t)partial_elems(rows, cols, true);
    if ((rc = c.ensure(&c.scratch_b, &c.scratch_b_bytes,
                       2 * qbytes + (size_t)(rows + 64) * sizeof(double) + npart * sizeof(double))))
        return finish(info, rc, t0);
    double *dR = (double *)c.scratch_a, *dQt = (double *)c.scratch_b, *dQ = dQt + (size_t)rows * pq;
    double *dv = dQ + (size_t)rows * pq;
    int *dflag = (int *)(dv + rows);
    double *dpart = dv + rows + 64;
    if ((rc = scirs::copy_h2d(dR, pr * sizeof(double), a, lda * sizeof(double), cols * sizeof(double), rows, s)))
        return finish(info, rc, t0);
    eye_kernel<<<(unsigned)(((long)rows * rows + 255) / 256), 256, 0, s>>>(dQt, pq, rows);
    if ((rc = householder_sweep(rows, cols, cols, dR, pr, dQt, pq, dv, dflag, dpart, s))) return finish(info, rc, t0);
    dim3 tg((rows + 31) / 32, (rows + 31) / 32);
    transpose_kernel<<<tg, 256, 0, s>>>(dQt, pq, rows, rows, dQ, pq);
    SCIRS_HIP_CHECK(hipGetLastError());
    if ((rc = scirs::copy_d2h(q, ldq * sizeof(double), dQ, pq * sizeof(double), rows * sizeof(double), rows, s)))
        return finish(info, rc, t0);
    if ((rc = scirs::copy_d2h(r, ldr * sizeof(double), dR, pr * sizeof(double), cols * sizeof(double), rows, s)))
        return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    return finish(info, SCIRS_OK, t0);
}

extern "C" int scirs_gpu_dlstsq(int rows, int cols, const double *a, int64_t lda, const double *b, double *x,
                                double *residual, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    if (rows <= 0 || cols <= 0 || rows < cols || !a || !b || !x || !residual || lda < cols) {
        scirs_set_error("dlstsq: invalid arguments (rows=%d cols=%d; rows >= cols >= 1)", rows, cols);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    Context &c = ctx();
    int rc = c.init(0);
    if (rc) return finish(info, rc, t0);
    hipStream_t s = c.stream;
    // R_ext = [A | b], rows x (cols + 1), padded to a 512-byte row pitch; b's
    // column takes every reflector, which forms Q^T b without Q.
    const long pr = round_up(cols + 1, 64), pa = round_up(cols, 64);
    const size_t rbytes = (size_t)rows * pr * sizeof(double);
    const size_t npart = (size_t)partial_elems(rows, cols + 1, false);
    const size_t wbytes = ((size_t)rows * pa + 3 * (size_t)rows + 2 * (size_t)cols + 128 + npart) * sizeof(double);
    if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, rbytes))) return finish(info, rc, t0);
    if ((rc = c.ensure(&c.scratch_b, &c.scratch_b_bytes, wbytes))) return finish(info, rc, t0);
    double *dR = (double *)c.scratch_a;
    // dy (the substitution's right-hand side and solution) starts 512-byte aligned
    double *dA = (double *)c.scratch_b, *dy = dA + (size_t)rows * pa, *db = dy + cols + 64, *dv = db + rows;
    double *dsq = dv + rows, *dres = dsq + rows;
    int *dflag = (int *)(dres + 1);
    double *dpart = dres + 64;
    if ((rc = scirs::copy_h2d(dR, pr * sizeof(double), a, lda * sizeof(double), cols * sizeof(double), rows, s)))
        return finish(info, rc, t0);
    if ((rc = scirs::copy_h2d(dA, pa * sizeof(double), a, lda * sizeof(double), cols * sizeof(double), rows, s)))
        return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipMemcpyAsync(db, b, rows * sizeof(double), hipMemcpyHostToDevice, s));
    // b into the extra column; strided columns move by kernels, only whole
    // contiguous vectors cross PCIe
    strided_scatter_kernel<<<(rows + 255) / 256, 256, 0, s>>>(db, rows, dR + cols, pr);
    SCIRS_HIP_CHECK(hipGetLastError());
    if ((rc = householder_sweep(rows, cols, cols + 1, dR, pr, nullptr, 0, dv, dflag, dpart, s)))
        return finish(info, rc, t0);
    // the back substitution's |r_ii| < eps test (solve.rs:184-191), first row
    // in substitution order (bottom up)
    std::vector<double> diag(cols);
    strided_gather_kernel<<<(cols + 255) / 256, 256, 0, s>>>(dR, pr + 1, cols, dsq);
    SCIRS_HIP_CHECK(hipGetLastError());
    SCIRS_HIP_CHECK(hipMemcpyAsync(diag.data(), dsq, cols * sizeof(double), hipMemcpyDeviceToHost, s));
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = cols - 1; i >= 0; --i)
        if (fabs(diag[i]) < DBL_EPSILON) {
            if (info) info->pivot_value = fabs(diag[i]);
            return finish(info, i + 1, t0);
        }
    strided_gather_kernel<<<(cols + 255) / 256, 256, 0, s>>>(dR + cols, pr, cols, dy);
    SCIRS_HIP_CHECK(hipGetLastError());
    if ((rc = c.ensure_xtri(cols))) return finish(info, rc, t0);
    if ((rc = c.reset_flags(s))) return finish(info, rc, t0);
    if ((rc = scirs_trsv_chain<double>(cols, 1, dR, pr, dy, 1, 0, 0, c.work, s))) return finish(info, rc, t0);
    int st = 0;
    if ((rc = c.read_flags(s, &st, nullptr))) return finish(info, rc, t0);
    resid_sq_kernel<<<(rows + 3) / 4, 256, 0, s>>>(dA, pa, rows, cols, db, dy, dsq);
    sum_kernel<<<1, HT, 0, s>>>(dsq, rows, dres);
    SCIRS_HIP_CHECK(hipGetLastError());
    SCIRS_HIP_CHECK(hipMemcpyAsync(x, dy, cols * sizeof(double), hipMemcpyDeviceToHost, s));
    SCIRS_HIP_CHECK(hipMemcpyAsync(residual, dres, sizeof(double), hipMemcpyDeviceToHost, s));
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    return finish(info, SCIRS_OK, t0);
}
