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
// of a row a 256-byte coalesced segment.  This unblocked (level-2) form is
// bound by passes over the trailing block per column; matrices of up to
// 16384 rows take the blocked (compact-WY) form further down, which applies
// 32 reflectors per pass (SCIRS_GPU_QR_BLOCKED=0 selects the unblocked one).
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

// ---------------------------------------------------------------------------
// Blocked form (compact WY, Schreiber & Van Loan 1989): the same reflectors,
// QB = 32 columns at a time.  H_1 ... H_w = I - V T V^T with V the unit
// reflectors as columns and T upper triangular (T_ii = 2 for an applied
// reflector, 0 for a skipped column), so the panel's reflectors reach the
// trailing columns of R and the rows k0.. of Q^T as
//     C <- C - V (T^T (V^T C))
// — one pass of partial products V^T C, one rank-32 product on the matrix
// cores — instead of two passes per column.
constexpr int QB = 32;        // panel width
constexpr int PR = 256;       // rows per panel workgroup, one per thread
constexpr int QG_MAX = 64;    // panel workgroups: rows - k0 <= QG_MAX * PR
constexpr int QKC = 32;       // row chunks of the partial products V^T C (at most)
constexpr int QWT = 256;      // columns per partial-product workgroup
typedef unsigned long long u64;
constexpr size_t QXW_WORDS = 2ull * QG_MAX * (QB + 1) * 2;   // exchange words (2 per value)

__device__ __forceinline__ void qx_put(u64 *w, unsigned seq, double v)
{
    const u64 bits = (u64)__double_as_longlong(v);
    __hip_atomic_store(w, ((u64)seq << 32) | (bits & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 1, ((u64)seq << 32) | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool qx_get(const u64 *w, unsigned seq, double &v)
{
    const u64 lo = __hip_atomic_load(const_cast<u64 *>(w), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const u64 hi = __hip_atomic_load(const_cast<u64 *>(w + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v = __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffull)));
    return (unsigned)(lo >> 32) == seq && (unsigned)(hi >> 32) == seq;
}
// poll one tagged value; false when the wait gave up (or another workgroup did)
__device__ bool qx_wait(const u64 *w, unsigned seq, double &v, unsigned spin_limit, int *abort)
{
    for (unsigned it = 0;; ++it) {
        if (qx_get(w, seq, v)) return true;
        if (it >= spin_limit || __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            __hip_atomic_store(abort, (int)SCIRS_ABORT_QR_PANEL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

struct QPanelArgs {
    double *R; long ldr;
    int rows, k0, w, G;
    double *V;          // rows x QB, ld QB: the panel's reflectors (0 above their first row, 0 past w)
    int *flags;         // [QB]: 1 reflector applied, 0 column skipped
    u64 *xw;            // [2][QG_MAX][QB + 1][2]: phase 0 (sums of squares, x0), phase 1 (dots)
    unsigned seq;       // phase p of column j uses seq + 2 j + p
    int *abort; unsigned spin_limit;
};

// The reflectors of columns k0 .. k0 + w - 1 with the reference's arithmetic
// (lapack.rs:205-235): per column the sum of squares, alpha = -sign(x0) ||x||,
// v = (x - alpha e1) / ||v||, and H = I - 2 v v^T on the panel's columns
// j.. .  G workgroups hold PR rows each in LDS; the two sums a column needs
// (||x||^2 and the dots v^T a_c) are exchanged as sequence-tagged words and
// added in workgroup order, so every workgroup computes the same alpha and v.
// The exchanged sum covers the rows below the pivot row, so ||x||^2 = x0^2 +
// tail and ||v||^2 = (x0 - alpha)^2 + tail (as the reference sums v_i^2, v0
// first) come from one exchange without cancellation.
__global__ __launch_bounds__(PR) void qr_panel_kernel(QPanelArgs a)
{
    // products rounded before they are added, as in the reference's loops
#pragma clang fp contract(off)
    __shared__ double P[PR][QB + 1];
    __shared__ double xd[QG_MAX][QB];
    __shared__ double red[PR / 32][QB];
    __shared__ double vcol[PR];
    __shared__ double tot[QB + 1];
    __shared__ int s_fail;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, g = blockIdx.x;
    const int row = a.k0 + g * PR + t;
    const bool have = row < a.rows;
    for (int c = 0; c < QB; ++c) P[t][c] = (have && c < a.w) ? a.R[(long)row * a.ldr + a.k0 + c] : 0.0;
    if (t == 0) s_fail = 0;
    __syncthreads();
    u64 *x0w = a.xw, *x1w = a.xw + (size_t)QG_MAX * (QB + 1) * 2;
    for (int j = 0; j < a.w; ++j) {
        const int rj = a.k0 + j;
        const unsigned sq0 = a.seq + 2u * j, sq1 = sq0 + 1u;
        // (a) this workgroup's sum of squares of column j below the pivot row
        //     (rows > rj); the pivot row rj lies in workgroup 0, which also
        //     publishes x0
        const double x = (have && row >= rj) ? P[t][j] : 0.0;
        const double ws = wave_sum(row > rj ? x * x : 0.0);
        if (lane == 0) red[wave][0] = ws;
        __syncthreads();
        if (t == 0) {
            double ss = 0.0;
            for (int q = 0; q < PR / 64; ++q) ss += red[q][0];
            qx_put(x0w + ((size_t)g * (QB + 1)) * 2, sq0, ss);
            qx_put(x0w + ((size_t)g * (QB + 1) + 1) * 2, sq0, g == 0 ? P[j][j] : 0.0);
        }
        if (wave == 0) {
            double v = 0.0;
            bool ok = true;
            if (lane < a.G) ok = qx_wait(x0w + ((size_t)lane * (QB + 1)) * 2, sq0, v, a.spin_limit, a.abort);
            double x0 = 0.0;
            if (lane == 0 && ok) ok = qx_wait(x0w + 2, sq0, x0, a.spin_limit, a.abort);
            const double sum = wave_sum(v);
            if (!ok) s_fail = 1;
            if (lane == 0) {
                tot[0] = sum;
                tot[1] = x0;
            }
        }
        __syncthreads();
        if (s_fail) return;
        const double tail = tot[0], x0 = tot[1];
        const double xn = sqrt(x0 * x0 + tail);
        bool apply = xn > DBL_EPSILON;
        double alpha = 0.0, v0 = 0.0, vn = 1.0;
        if (apply) {
            alpha = x0 >= 0.0 ? -xn : xn;
            v0 = x0 - alpha;
            vn = sqrt(v0 * v0 + tail);
            apply = vn > DBL_EPSILON;
        }
        if (g == 0 && t == 0) a.flags[j] = apply ? 1 : 0;
        const double vi = (!apply || !have || row < rj) ? 0.0 : (row == rj ? v0 / vn : x / vn);
        if (have) a.V[(long)row * QB + j] = vi;
        if (!apply) continue;   // the same decision in every workgroup
        vcol[t] = vi;
        __syncthreads();
        // (b) dots d_c = v^T a_c for the panel's columns c >= j: 32 columns x
        //     8 row groups, the groups added in order, then exchanged
        {
            const int cc = t & (QB - 1), rg = t >> 5;
            double d = 0.0;
            if (cc >= j && cc < a.w)
                for (int i = rg; i < PR; i += PR / 32) d += vcol[i] * P[i][cc];
            red[rg][cc] = d;
        }
        __syncthreads();
        if (t < QB && t >= j && t < a.w) {
            double d = 0.0;
            for (int q = 0; q < PR / 32; ++q) d += red[q][t];
            qx_put(x1w + ((size_t)g * (QB + 1) + t) * 2, sq1, d);
        }
        for (int q = t; q < a.G * QB; q += PR) {
            const int gq = q / QB, cq = q % QB;
            if (cq >= j && cq < a.w) {
                double v;
                if (!qx_wait(x1w + ((size_t)gq * (QB + 1) + cq) * 2, sq1, v, a.spin_limit, a.abort)) s_fail = 1;
                xd[gq][cq] = v;
            }
        }
        __syncthreads();
        if (s_fail) return;
        if (t < QB && t >= j && t < a.w) {
            double d = 0.0;
            for (int q = 0; q < a.G; ++q) d += xd[q][t];
            tot[t] = d;
        }
        __syncthreads();
        // (c) H on this workgroup's rows of columns j.. (the reference's
        //     r_ic -= 2 v_i (v^T c))
        if (have)
            for (int c = j; c < a.w; ++c) P[t][c] -= 2.0 * vi * tot[c];
        __syncthreads();
    }
    if (have) {
        for (int c = 0; c < a.w; ++c) a.R[(long)row * a.ldr + a.k0 + c] = P[t][c];
        for (int c = a.w; c < QB; ++c) a.V[(long)row * QB + c] = 0.0;
    }
}

// Partial products over row chunks: part[kc][jj][col] = sum over the chunk's
// rows of V[row][jj] * C[row][col], for the virtual columns
//   [0, QB)             the panel's own reflectors (V^T V, for T),
//   [QB, QB + nR)       R's columns rc0 .. rc0 + nR - 1,
//   [QB + nR, ntot)     Q^T's columns 0 .. nQ - 1.
// One thread per column, every row of the chunk; V's rows staged in LDS and
// read as broadcasts.
struct QWArgs {
    const double *V;
    const double *R; long ldr; int rc0, nR;
    const double *Qt; long ldq; int nQ;
    int rows, k0, rch;
    double *part; long ldp;
};
__global__ __launch_bounds__(QWT) void qr_wpart_kernel(QWArgs a)
{
    // pitch QB: every read is a broadcast, and pairs are 16-byte aligned
    __shared__ double Vs[64][QB];
    const int t = threadIdx.x, col = blockIdx.x * QWT + t, kc = blockIdx.y;
    const int ntot = QB + a.nR + a.nQ;
    const double *src = nullptr;
    long ld = 0;
    if (col < QB) { src = a.V + col; ld = QB; }
    else if (col < QB + a.nR) { src = a.R + a.rc0 + (col - QB); ld = a.ldr; }
    else if (col < ntot) { src = a.Qt + (col - QB - a.nR); ld = a.ldq; }
    double acc[QB];
#pragma unroll
    for (int jj = 0; jj < QB; ++jj) acc[jj] = 0.0;
    const int rbeg = a.k0 + kc * a.rch, rend = min(a.rows, rbeg + a.rch);
    for (int r0 = rbeg; r0 < rend; r0 += 64) {
        __syncthreads();
        for (int e = t; e < 64 * QB; e += QWT) {
            const int rr = e / QB, jj = e % QB;
            Vs[rr][jj] = r0 + rr < rend ? a.V[(long)(r0 + rr) * QB + jj] : 0.0;
        }
        __syncthreads();
        const int nrow = min(64, rend - r0);
        if (src)
            // sixteen rows' loads in flight before their products (rows past
            // the chunk read row r0 again and are multiplied by zero rows of Vs)
            for (int rb = 0; rb < nrow; rb += 16) {
                double cv[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) cv[u] = src[(long)(r0 + min(rb + u, nrow - 1)) * ld];
#pragma unroll
                for (int u = 0; u < 16; ++u)
#pragma unroll
                    for (int jj = 0; jj < QB; ++jj) acc[jj] = fma(Vs[rb + u][jj], cv[u], acc[jj]);
            }
    }
    if (col < ntot)
#pragma unroll
        for (int jj = 0; jj < QB; ++jj) a.part[((long)kc * QB + jj) * a.ldp + col] = acc[jj];
}

// T (QB x QB, upper) from the Gram matrix V^T V (the partials' first QB
// columns, added in chunk order): column i is -2 T[0:i, 0:i] (V^T v_i) above
// T_ii = 2 (LAPACK's forward column-wise T); a skipped column is all zero.
// Row t of T needs only its own earlier entries and the Gram matrix, so lane
// t of one wave builds it with no barrier.
__global__ __launch_bounds__(256) void qr_tbuild_kernel(const double *part, long ldp, int kcn, const int *flags,
                                                       double *T)
{
    __shared__ double Gm[QB][QB + 1];
    __shared__ double Ts[QB][QB + 1];
    const int t = threadIdx.x;
    for (int e = t; e < QB * QB; e += 256) {
        const int r = e / QB, c = e % QB;
        double sacc = 0.0;
#pragma unroll 8
        for (int kc = 0; kc < kcn; ++kc) sacc += part[((long)kc * QB + r) * ldp + c];
        Gm[r][c] = sacc;
    }
    __syncthreads();
    if (t < QB) {
        // the row in registers; the Gram entries are broadcast reads
        double tr[QB];
#pragma unroll
        for (int i = 0; i < QB; ++i) {
            const double tau = flags[i] != 0 ? 2.0 : 0.0;
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < i; ++q) v += (q >= t ? tr[q] : 0.0) * (-tau * Gm[q][i]);
            tr[i] = t < i ? v : (t == i ? tau : 0.0);
        }
#pragma unroll
        for (int i = 0; i < QB; ++i) Ts[t][i] = tr[i];
    }
    __syncthreads();
    for (int e = t; e < QB * QB; e += 256) T[e] = Ts[e / QB][e % QB];
}

// Y[:, col] = T^T W[:, col], W = the partials added in chunk order, for the
// virtual columns QB .. ntot-1 (Y is QB x (ntot - QB), ld ldy).  A workgroup
// takes 8 columns: thread (jj, col) adds the kcn partials of W[jj][col] (all
// its loads in flight at once), then thread (i, col) forms row i of Y.
constexpr int QYC = 256 / QB;   // columns per workgroup
__global__ __launch_bounds__(256) void qr_y_kernel(const double *part, long ldp, int kcn, int ntot, const double *T,
                                                   double *Y, long ldy)
{
    __shared__ double Ts[QB][QB + 1];
    __shared__ double Wl[QB][QYC];
    const int t = threadIdx.x, jj = t / QYC, cl = t % QYC;
    for (int e = t; e < QB * QB; e += 256) Ts[e / QB][e % QB] = T[e];
    const int col = QB + blockIdx.x * QYC + cl;
    double w = 0.0;
    if (col < ntot) {
#pragma unroll 8
        for (int kc = 0; kc < kcn; ++kc) w += part[((long)kc * QB + jj) * ldp + col];
    }
    Wl[jj][cl] = w;
    __syncthreads();
    if (col < ntot) {
        const int i = jj;
        double y = 0.0;
        for (int q = 0; q <= i; ++q) y = fma(Ts[q][i], Wl[q][cl], y);
        Y[(long)i * ldy + (col - QB)] = y;
    }
}

__global__ void fill_u64_kernel(u64 *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// The blocked sweep's device workspace (grown on demand; the exchange words
// keep their tags across calls, so the sequence continues).
struct QrBlockedWork {
    double *V = nullptr, *part = nullptr, *T = nullptr, *Y = nullptr;
    int *flags = nullptr;
    u64 *xw = nullptr;
    size_t bytes = 0;
    void *base = nullptr;
    unsigned seq = 1;
};
QrBlockedWork g_qbw;

bool blocked_ok(int rows)
{
    static const int on = [] {
        const char *v = getenv("SCIRS_GPU_QR_BLOCKED");
        return !(v && *v && atoi(v) == 0);
    }();
    return on && rows <= QG_MAX * PR;
}

int householder_blocked(int rows, int cols, int rcols, double *R, long ldr, double *Qt, long ldq, hipStream_t s)
{
    Context &c = ctx();
    const int kmax = std::min(rows, cols);
    const int nQ = Qt ? rows : 0;
    const long ldp = round_up(QB + rcols + nQ, 64);
    const size_t need = ((size_t)rows * QB + (size_t)QKC * QB * ldp + QB * QB + (size_t)QB * ldp + 64) * sizeof(double) +
                        QXW_WORDS * sizeof(u64) + 256;
    QrBlockedWork &q = g_qbw;
    if (need > q.bytes) {
        if (q.base) {
            SCIRS_HIP_CHECK(hipDeviceSynchronize());
            SCIRS_HIP_CHECK(hipFree(q.base));
            q.base = nullptr;
            q.bytes = 0;
        }
        if (hipMalloc(&q.base, need) != hipSuccess) {
            q.base = nullptr;
            scirs_set_error("blocked QR: device allocation of %zu bytes failed", need);
            return SCIRS_ERR_NOMEM;
        }
        q.bytes = need;
        q.xw = (u64 *)q.base;
        // tag 0 is never a live sequence number
        fill_u64_kernel<<<(unsigned)((QXW_WORDS + 255) / 256), 256, 0, s>>>(q.xw, QXW_WORDS);
        SCIRS_HIP_CHECK(hipGetLastError());
        q.seq = 1;
    }
    // this call's layout (the buffer may be larger than it needs)
    q.V = (double *)(q.xw + QXW_WORDS);
    q.part = q.V + (size_t)rows * QB;
    q.T = q.part + (size_t)QKC * QB * ldp;
    q.Y = q.T + QB * QB;
    q.flags = (int *)(q.Y + (size_t)QB * ldp);
    for (int k0 = 0; k0 < kmax; k0 += QB) {
        const int w = std::min(QB, kmax - k0), rows_k = rows - k0;
        if (q.seq > 0xFFFFFFFFu - 2u * QB) {
            fill_u64_kernel<<<(unsigned)((QXW_WORDS + 255) / 256), 256, 0, s>>>(q.xw, QXW_WORDS);
            q.seq = 1;
        }
        QPanelArgs pa;
        pa.R = R; pa.ldr = ldr; pa.rows = rows; pa.k0 = k0; pa.w = w;
        pa.G = (rows_k + PR - 1) / PR;
        pa.V = q.V; pa.flags = q.flags; pa.xw = q.xw; pa.seq = q.seq;
        pa.abort = c.work.abort; pa.spin_limit = c.work.spin_limit;
        q.seq += 2u * QB;
        qr_panel_kernel<<<pa.G, PR, 0, s>>>(pa);
        const int rc0 = k0 + w, nR = rcols - rc0, ntot = QB + nR + nQ;
        // enough row chunks for about 512 workgroups, no more (every chunk adds
        // a partial that T and Y read back)
        const int ncb = (ntot + QWT - 1) / QWT;
        int kcn = std::min(std::min(QKC, (rows_k + 63) / 64), std::max(1, (512 + ncb - 1) / ncb));
        const int rch = (int)round_up((rows_k + kcn - 1) / kcn, 64);
        kcn = (rows_k + rch - 1) / rch;
        QWArgs wa;
        wa.V = q.V; wa.R = R; wa.ldr = ldr; wa.rc0 = rc0; wa.nR = nR; wa.Qt = Qt; wa.ldq = ldq; wa.nQ = nQ;
        wa.rows = rows; wa.k0 = k0; wa.rch = rch; wa.part = q.part; wa.ldp = ldp;
        qr_wpart_kernel<<<dim3((ntot + QWT - 1) / QWT, kcn), QWT, 0, s>>>(wa);
        qr_tbuild_kernel<<<1, 256, 0, s>>>(q.part, ldp, kcn, q.flags, q.T);
        if (nR + nQ > 0) {
            qr_y_kernel<<<(nR + nQ + QYC - 1) / QYC, 256, 0, s>>>(q.part, ldp, kcn, ntot, q.T, q.Y, ldp);
            const double *Vk = q.V + (size_t)k0 * QB;
            if (nR > 0)
                if (int rc = scirs_launch_dgemm_sub(rows_k, nR, QB, Vk, QB, q.Y, ldp, 0, R + (long)k0 * ldr + rc0, ldr,
                                                    0, c.work.info, s))
                    return rc;
            if (nQ > 0)
                if (int rc = scirs_launch_dgemm_sub(rows_k, nQ, QB, Vk, QB, q.Y + nR, ldp, 0, Qt + (long)k0 * ldq,
                                                    ldq, 0, c.work.info, s))
                    return rc;
        }
        SCIRS_HIP_CHECK(hipGetLastError());
    }
    return SCIRS_OK;
}

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
    if ((rc = c.reset_flags(s))) return finish(info, rc, t0);
    if ((rc = blocked_ok(rows) ? householder_blocked(rows, cols, cols, dR, pr, dQt, pq, s)
                               : householder_sweep(rows, cols, cols, dR, pr, dQt, pq, dv, dflag, dpart, s)))
        return finish(info, rc, t0);
    {
        int st = 0;
        if ((rc = c.read_flags(s, &st, nullptr))) return finish(info, rc, t0);
    }
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
    if ((rc = c.reset_flags(s))) return finish(info, rc, t0);
    if ((rc = blocked_ok(rows) ? householder_blocked(rows, cols, cols + 1, dR, pr, nullptr, 0, s)
                               : householder_sweep(rows, cols, cols + 1, dR, pr, nullptr, 0, dv, dflag, dpart, s)))
        return finish(info, rc, t0);
    {
        int st0 = 0;
        if ((rc = c.read_flags(s, &st0, nullptr))) return finish(info, rc, t0);
    }
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
