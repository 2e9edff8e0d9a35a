// extern "C" entry points of include/scirs_gpu.h: one process-wide context per
// device (stream, scratch, workspace), host<->HBM staging for the host-memory
// entry points, and the plain FP64 / FP32 factor-and-solve drivers.
#include "common.hpp"
#include "kernels.hpp"
#include "api_internal.hpp"
#include "../../include/scirs_gpu.h"

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {
thread_local std::string g_last_error;
std::mutex g_err_mu;
std::string g_last_error_any;
} // namespace

extern "C" void scirs_set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_error_any = buf;
    if (getenv("SCIRS_GPU_LOG")) fprintf(stderr, "[scirs_gpu] %s\n", buf);
}

int scirs_allow_lds(const void *kernel, size_t dynamic_bytes)
{
    static std::mutex mu;
    static std::vector<std::pair<const void *, size_t>> granted;
    std::lock_guard<std::mutex> lk(mu);
    for (auto &g : granted)
        if (g.first == kernel && g.second >= dynamic_bytes) return SCIRS_OK;
    hipFuncAttributes attr;
    SCIRS_HIP_CHECK(hipFuncGetAttributes(&attr, kernel));
    const size_t limit = 160 * 1024 - attr.sharedSizeBytes;
    if (dynamic_bytes > limit) {
        scirs_set_error("kernel needs %zu B of dynamic LDS; only %zu B are available", dynamic_bytes, limit);
        return SCIRS_ERR_HIP;
    }
    SCIRS_HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)limit));
    for (auto &g : granted)
        if (g.first == kernel) { g.second = limit; return SCIRS_OK; }
    granted.emplace_back(kernel, limit);
    return SCIRS_OK;
}

extern "C" const char *scirs_gpu_last_error(void)
{
    if (!g_last_error.empty()) return g_last_error.c_str();
    return "";
}

// Scratch of scirs_gpu_backward_error_dev_noalloc: a fixed array in the code
// object, so that entry point calls no allocator.  Host code must use the
// address hipGetSymbolAddress returns for it (Context::noalloc_scratch), never
// the array's name, which on the host is a stand-in and not a device address.
#define SCIRS_NOALLOC_SCRATCH_DOUBLES (1u << 20)   /* 8 MiB */
static __device__ __align__(256) double scratch[SCIRS_NOALLOC_SCRATCH_DOUBLES];

namespace scirs {

static Context g_ctx;

Context &ctx() { return g_ctx; }

static int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

int Context::ensure(void **p, size_t *have, size_t bytes)
{
    if (*have >= bytes && *p) return SCIRS_OK;
    if (*p) {
        // the old buffer may still be in use on a caller's stream
        SCIRS_HIP_CHECK(hipDeviceSynchronize());
        SCIRS_HIP_CHECK(hipFree(*p));
        *p = nullptr;
        *have = 0;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        scirs_set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        *p = nullptr;
        return SCIRS_ERR_NOMEM;
    }
    *have = bytes;
    return SCIRS_OK;
}

int Context::init(int dev)
{
    if (initialised) return SCIRS_OK;
    int count = 0;
    SCIRS_HIP_CHECK(hipGetDeviceCount(&count));
    if (count <= 0) {
        scirs_set_error("no HIP device visible");
        return SCIRS_ERR_HIP;
    }
    if (dev < 0 || dev >= count) {
        scirs_set_error("device %d out of range (%d devices)", dev, count);
        return SCIRS_ERR_ARG;
    }
    device = dev;
    SCIRS_HIP_CHECK(hipSetDevice(dev));
    hipDeviceProp_t prop;
    SCIRS_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    cu_count = prop.multiProcessorCount;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        scirs_set_error("device %d is %s; this library is built for gfx950 only", dev, prop.gcnArchName);
        return SCIRS_ERR_UNSUPPORTED;
    }
    SCIRS_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // side stream of the look-ahead LU (default priority, like every stream here)
    SCIRS_HIP_CHECK(hipStreamCreateWithFlags(&work.side_stream, hipStreamNonBlocking));
    SCIRS_HIP_CHECK(hipEventCreateWithFlags(&work.ev_trsm, hipEventDisableTiming));
    SCIRS_HIP_CHECK(hipEventCreateWithFlags(&work.ev_upd, hipEventDisableTiming));
    SCIRS_HIP_CHECK(hipEventCreateWithFlags(&work.ev_start, hipEventDisableTiming));
    // factorisation workspace
    const int max_g = SCIRS_MAX_G;
    const int kbmax = SCIRS_KB_MAX;
    work.max_g = std::min(env_int("SCIRS_GPU_MAX_G", cu_count), max_g);
    work.max_panels = 1 << 14;   // n up to 16384 * kb
    work.spin_limit = (unsigned)env_int("SCIRS_GPU_SPIN_LIMIT", 1 << 22);
    work.lowp_mode = SCIRS_LOWP_FP32;   // plain FP32 entry points compute in FP32; the mixed solve switches it
    work.panel_rows = env_int("SCIRS_GPU_PANEL_ROWS", 192);
    work.panel_w0 = env_int("SCIRS_GPU_PANEL_W0", 0);
    work.lookahead = env_int("SCIRS_GPU_LOOKAHEAD", 1);
    work.trsv_pair = env_int("SCIRS_GPU_TRSV_PAIR", 1);
    work.poll_sleep = env_int("SCIRS_GPU_POLL_SLEEP", 0);
    work.panel_nrep = std::max(1, std::min((int)SCIRS_PANEL_REP_MAX, env_int("SCIRS_GPU_PANEL_REPLICAS", 1)));
    size_t bytes = 0;
    const size_t rep = SCIRS_PANEL_REP_MAX;
    const size_t xw_b = rep * 2 * (size_t)max_g * 4 * 8, xrow_b = rep * 2 * (size_t)max_g * kbmax * 2 * 8,
                 xrj_b = rep * 2 * (size_t)kbmax * 2 * 8, flags_b = 64 * sizeof(int),
                 pairs_b = (size_t)work.max_panels * 4 * kbmax * sizeof(int),
                 npairs_b = (size_t)work.max_panels * sizeof(int),
                 linv_b = 2 * 128 * 128 * sizeof(float),   // two FP32 L11^-1 (double-buffered)
                 linv64_b = 128 * 128 * sizeof(double);
    const size_t tpc_b = (size_t)TSLU_MAXLEV * TSLU_MAXNODE * 65 * 8, tpn_b = (size_t)TSLU_MAXLEV * TSLU_MAXNODE * 4,
                 tpu_b = 64 * 64 * sizeof(float), tpr_b = 64 * 64 * sizeof(float),
                 tpt_b = 64 * (size_t)kbmax * sizeof(float), tpv_b = 65 * sizeof(int),
                 tpm_b = (2 * (size_t)TSLU_MAP + 1) * sizeof(int);
    for (size_t b : {xw_b, xrow_b, xrj_b, flags_b, pairs_b, npairs_b, linv_b, linv64_b, tpc_b, tpn_b, tpu_b, tpr_b, tpt_b, tpt_b,
                     tpv_b, tpm_b})
        bytes += (b + 255) & ~(size_t)255;
    char *base = nullptr;
    SCIRS_HIP_CHECK(hipMalloc((void **)&base, bytes));
    SCIRS_HIP_CHECK(hipMemset(base, 0, bytes));
    char *p = base;
    auto take = [&](size_t n) { char *r = p; p += (n + 255) & ~(size_t)255; return r; };
    work.xw = (unsigned long long *)take(xw_b);
    work.xrow = (unsigned long long *)take(xrow_b);
    work.xrj = (unsigned long long *)take(xrj_b);
    work.xw_words = xw_b / 8;
    work.xrow_words = xrow_b / 8;
    work.xrj_words = xrj_b / 8;
    char *flags = take(flags_b);
    work.abort = (int *)(flags + 8);
    work.info = (int *)(flags + 12);
    work.info_val = (double *)(flags + 16);
    work.ticket = (int *)(flags + 32);
    work.pairs = (int *)take(pairs_b);
    work.npairs = (int *)take(npairs_b);
    work.linv = take(linv_b);
    work.linv64 = (double *)take(linv64_b);
    work.tp_cand = (unsigned long long *)take(tpc_b);
    work.tp_cnt = (int *)take(tpn_b);
    work.tp_uinv = (float *)take(tpu_b);
    work.tp_urest = (float *)take(tpr_b);
    work.tp_top = (float *)take(tpt_b);
    work.tp_stage = (float *)take(tpt_b);
    work.tp_vlist = (int *)take(tpv_b);
    work.tp_map = (int *)take(tpm_b);
    work.tslu = env_int("SCIRS_GPU_TSLU", 1);
    work.seq_next = 1;
    work_base = base;
    SCIRS_HIP_CHECK(hipHostMalloc((void **)&host_flags, 64, hipHostMallocDefault));
    // this device's copy of the allocation-free entry point's scratch
    SCIRS_HIP_CHECK(hipGetSymbolAddress((void **)&noalloc_scratch, HIP_SYMBOL(scratch)));
    // panel widths: 8 .. 128 (the LDS budgets of the panel's neighbours — the
    // next-block interchange + TRSM kernel, the TRSM and Cholesky kernels — are
    // sized for kb <= 128)
    auto kb_env = [](const char *name) { return std::max(8, std::min(128, env_int(name, 128))); };
    kb_f64 = kb_env("SCIRS_GPU_KB");
    kb_f32 = kb_env("SCIRS_GPU_KB_F32");
    kb_chol = kb_env("SCIRS_GPU_KB_CHOL");
    narrow_kb = env_int("SCIRS_GPU_NARROW_KB", 1);
    chol_chain = env_int("SCIRS_GPU_CHOL_CHAIN", 1);
    trsv_dinv = env_int("SCIRS_GPU_TRSV_DINV", 1);
    initialised = true;
    return SCIRS_OK;
}

// Reset info / abort before a factorisation (a kernel, so it is safe under capture).
int Context::reset_flags(hipStream_t s)
{
    return scirs_fill_int(work.abort, 2, 0, s);   // abort and info are adjacent
}

// Read info, info_val and abort after the stream drains.
int Context::read_flags(hipStream_t s, int *info, double *info_val)
{
    SCIRS_HIP_CHECK(hipMemcpyAsync(host_flags, work.abort, 8 + 8, hipMemcpyDeviceToHost, s));
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    int abort_flag = ((int *)host_flags)[0];
    *info = ((int *)host_flags)[1];
    double v;
    memcpy(&v, (char *)host_flags + 8, 8);
    if (info_val) *info_val = v;
    if (abort_flag) {
        if (abort_flag > 0)
            scirs_set_error("the LU panel's exchange timed out at column %d (another process may hold the device's "
                            "compute units; SCIRS_GPU_SPIN_LIMIT bounds the wait)", abort_flag);
        else if (abort_flag == SCIRS_ABORT_QR_PANEL)
            scirs_set_error("the QR panel's exchange timed out (another process may hold the device's compute "
                            "units; SCIRS_GPU_SPIN_LIMIT bounds the wait)");
        else
            scirs_set_error("a chained triangular solve timed out waiting for its hand-over (SCIRS_GPU_SPIN_LIMIT "
                            "bounds the wait)");
        return SCIRS_ERR_TIMEOUT;
    }
    return SCIRS_OK;
}

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

void info_init(scirs_solve_info *info, int path)
{
    if (!info) return;
    memset(info, 0, sizeof(*info));
    info->path = path;
    info->backward_error = -1.0;
}

int finish(scirs_solve_info *info, int rc, double t0)
{
    if (info) {
        info->status = rc;
        info->time_total_ms = now_ms() - t0;
    }
    return rc;
}

// normwise backward error  max_j ||R_j||_inf / (||A||_inf ||X_j||_inf + ||B_j||_inf)
// Doubles of scratch the backward error uses: R (n x nrhs), row sums (n), norms.
static size_t backward_error_scratch(int n, int nrhs)
{
    return (size_t)n * nrhs + (size_t)n + 3 * 64 + 8;
}

// The backward error with its scratch (backward_error_scratch(n, nrhs) doubles)
// supplied.  No part of it needs to start at zero: every value read back is
// written by a kernel of this call first.
static int backward_error_in(int n, int nrhs, const double *dA, long lda, const double *dX, long ldx,
                             const double *dB, long ldb, hipStream_t s, double *scratch, double *out)
{
    int rc;
    double *R = scratch;
    double *rowsum = R + (size_t)n * nrhs;
    double *norms = rowsum + n;   // [0..63] r, [64..127] x, [128..191] b, [192] A
    if ((rc = scirs_residual(n, nrhs, dA, lda, dX, ldx, dB, ldb, R, nrhs, s))) return rc;
    if ((rc = scirs_convert_rowsum(n, n, dA, lda, nullptr, 0, rowsum, s))) return rc;
    if ((rc = scirs_rowsum_max(n, rowsum, norms + 192, s))) return rc;
    double worst = 0.0;
    for (int j0 = 0; j0 < nrhs; j0 += 64) {
        int nj = std::min(64, nrhs - j0);
        if ((rc = scirs_colmax_d(n, nj, R + j0, nrhs, norms, s))) return rc;
        if ((rc = scirs_colmax_d(n, nj, dX + j0, ldx, norms + 64, s))) return rc;
        if ((rc = scirs_colmax_d(n, nj, dB + j0, ldb, norms + 128, s))) return rc;
        double h[193];
        SCIRS_HIP_CHECK(hipMemcpyAsync(h, norms, 193 * sizeof(double), hipMemcpyDeviceToHost, s));
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
        for (int j = 0; j < nj; ++j) {
            double den = h[192] * h[64 + j] + h[128 + j];
            double e = den > 0 ? h[j] / den : (h[j] > 0 ? INFINITY : 0.0);
            worst = std::max(worst, e);
        }
    }
    *out = worst;
    return SCIRS_OK;
}

int backward_error(int n, int nrhs, const double *dA, long lda, const double *dX, long ldx, const double *dB,
                   long ldb, hipStream_t s, double *out)
{
    Context &c = ctx();
    if (int rc = c.ensure(&c.scratch_bwd, &c.scratch_bwd_bytes, backward_error_scratch(n, nrhs) * sizeof(double)))
        return rc;
    return backward_error_in(n, nrhs, dA, lda, dX, ldx, dB, ldb, s, (double *)c.scratch_bwd, out);
}

} // namespace scirs

using namespace scirs;

#define SCIRS_ENTER(dev_needed)                                   \
    std::lock_guard<std::recursive_mutex> _lk(ctx().mu);          \
    {                                                             \
        int _rc = ctx().init(env_int_dev());                      \
        if (_rc) return _rc;                                      \
        SCIRS_HIP_CHECK(hipSetDevice(ctx().device));              \
    }

static int env_int_dev()
{
    const char *v = getenv("SCIRS_GPU_DEVICE");
    return v && *v ? atoi(v) : 0;
}

extern "C" int scirs_gpu_init(int device)
{
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    return ctx().init(device < 0 ? 0 : device);
}

int scirs::Context::ensure_xtri(int n)
{
    const size_t words = (size_t)std::max(n, 1) * 4 * 2;
    if (words <= work.xtri_words && xtri_mem) return SCIRS_OK;
    int rc = ensure(&xtri_mem, &xtri_bytes, words * sizeof(unsigned long long));
    if (rc) return rc;
    // tag 0 is never a live sequence number, so zeroed words never match
    SCIRS_HIP_CHECK(hipMemset(xtri_mem, 0, words * sizeof(unsigned long long)));
    work.xtri = (unsigned long long *)xtri_mem;
    work.xtri_words = words;
    return SCIRS_OK;
}

// ------------------------------------------------------------------ LU core
// X <- U^-1 L^-1 X (X already row-permuted): the chained one-launch solver by
// default; SCIRS_GPU_TRSV=big selects the 256-row two-launches-per-block
// solver, SCIRS_GPU_TRSV=blocked the 64-row one.
template <typename T>
int scirs::lu_substitute(int n, int nrhs, const T *dLU, long ldlu, T *dX, long ldx, hipStream_t s, const T *dinv_l,
                         const T *dinv_u, const T *dmt_l, const T *dmt_u)
{
    Context &c = ctx();
    const char *mode = getenv("SCIRS_GPU_TRSV");
    if (mode && strcmp(mode, "blocked") == 0)
        return scirs_getrs_device<T>(n, nrhs, dLU, ldlu, dX, ldx, c.work.info, s);
    int rc;
    if (mode && strcmp(mode, "big") == 0) {
        rc = scirs_trsv_big<T>(n, nrhs, dLU, ldlu, dX, ldx, 1, 1, c.work.info, s);
        if (rc) return rc;
        return scirs_trsv_big<T>(n, nrhs, dLU, ldlu, dX, ldx, 0, 0, c.work.info, s);
    }
    if ((rc = c.ensure_xtri(n))) return rc;
    rc = scirs_trsv_chain<T>(n, nrhs, dLU, ldlu, dX, ldx, 1, 1, c.work, s, dinv_l, dmt_l);
    if (rc) return rc;
    return scirs_trsv_chain<T>(n, nrhs, dLU, ldlu, dX, ldx, 0, 0, c.work, s, dinv_u, dmt_u);
}
template int scirs::lu_substitute<float>(int, int, const float *, long, float *, long, hipStream_t, const float *,
                                         const float *, const float *, const float *);
template int scirs::lu_substitute<double>(int, int, const double *, long, double *, long, hipStream_t,
                                          const double *, const double *, const double *, const double *);

template <typename T>
static int getrf_solve_device(int n, int nrhs, T *dLU, long ldlu, int64_t *dPerm, const T *dB, long ldb, T *dX,
                              long ldx, hipStream_t s, scirs_solve_info *info)
{
    Context &c = ctx();
    int rc;
    scirs::CallEvents ev;
    if (int erc = ev.create(3)) return erc;
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1], e2 = ev.e[2];
    SCIRS_HIP_CHECK(hipEventRecord(e0, s));
    if ((rc = c.reset_flags(s))) return rc;
    const int kb = c.lu_kb<T>(n);
    if ((rc = scirs_getrf_device<T>(n, n, dLU, ldlu, kb, c.work, s))) return rc;
    if ((rc = scirs_perm_from_pairs(n, n, kb, c.work, dPerm, s))) return rc;
    SCIRS_HIP_CHECK(hipEventRecord(e1, s));
    if (nrhs > 0) {
        if ((rc = scirs_permute_rows<T>(n, nrhs, dPerm, dB, ldb, dX, ldx, s))) return rc;
        if ((rc = lu_substitute<T>(n, nrhs, dLU, ldlu, dX, ldx, s))) return rc;
    }
    SCIRS_HIP_CHECK(hipEventRecord(e2, s));
    int st = 0;
    double pv = 0;
    if ((rc = c.read_flags(s, &st, &pv))) return rc;
    if (info) {
        float ms1 = 0, ms2 = 0;
        (void)hipEventElapsedTime(&ms1, e0, e1);
        (void)hipEventElapsedTime(&ms2, e1, e2);
        info->time_factor_ms = ms1;
        info->time_solve_ms = ms2;
        info->pivot_value = st ? pv : 0.0;
    }
    return st;   // 0 or the 1-based singular column
}

template <typename T>
static int host_getrf(int m, int n, T *a, int64_t lda, int64_t *perm, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, sizeof(T) == 8 ? SCIRS_PATH_FP64_LU : SCIRS_PATH_FP32_LU);
    if (m <= 0 || n <= 0 || !a || !perm || lda < n) {
        scirs_set_error("getrf: invalid arguments (m=%d n=%d lda=%lld)", m, n, (long long)lda);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    SCIRS_ENTER(true);
    Context &c = ctx();
    hipStream_t s = c.stream;
    int rc;
    const size_t abytes = (size_t)m * n * sizeof(T);
    if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, abytes))) return finish(info, rc, t0);
    if ((rc = c.ensure(&c.scratch_perm, &c.scratch_perm_bytes, (size_t)m * sizeof(int64_t)))) return finish(info, rc, t0);
    T *dA = (T *)c.scratch_a;
    int64_t *dP = (int64_t *)c.scratch_perm;
    if ((rc = scirs::copy_h2d(dA, n * sizeof(T), a, lda * sizeof(T), n * sizeof(T), m, s))) return finish(info, rc, t0);
    if ((rc = c.reset_flags(s))) return finish(info, rc, t0);
    const int kb = c.lu_kb<T>(m);
    if ((rc = scirs_getrf_device<T>(m, n, dA, n, kb, c.work, s))) return finish(info, rc, t0);
    if ((rc = scirs_perm_from_pairs(m, n, kb, c.work, dP, s))) return finish(info, rc, t0);
    int st = 0;
    double pv = 0;
    if ((rc = c.read_flags(s, &st, &pv))) return finish(info, rc, t0);
    if (info) info->pivot_value = st ? pv : 0.0;
    if (st == 0) {
        if ((rc = scirs::copy_d2h(a, lda * sizeof(T), dA, n * sizeof(T), n * sizeof(T), m, s)))
            return finish(info, rc, t0);
        SCIRS_HIP_CHECK(hipMemcpyAsync(perm, dP, (size_t)m * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    }
    return finish(info, st, t0);
}

extern "C" int scirs_gpu_dgetrf(int m, int n, double *a, int64_t lda, int64_t *perm, scirs_solve_info *info)
{
    return host_getrf<double>(m, n, a, lda, perm, info);
}

extern "C" int scirs_gpu_sgetrf(int m, int n, float *a, int64_t lda, int64_t *perm, scirs_solve_info *info)
{
    return host_getrf<float>(m, n, a, lda, perm, info);
}

template <typename T>
static int host_gesv(int n, int nrhs, const T *a, int64_t lda, const T *b, int64_t ldb, T *x, int64_t ldx,
                     scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, sizeof(T) == 8 ? SCIRS_PATH_FP64_LU : SCIRS_PATH_FP32_LU);
    if (n <= 0 || nrhs <= 0 || !a || !b || !x || lda < n || ldb < nrhs || ldx < nrhs) {
        scirs_set_error("gesv: invalid arguments (n=%d nrhs=%d)", n, nrhs);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    SCIRS_ENTER(true);
    Context &c = ctx();
    hipStream_t s = c.stream;
    int rc;
    const size_t abytes = (size_t)n * n * sizeof(T), bbytes = (size_t)n * nrhs * sizeof(T);
    if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, abytes))) return finish(info, rc, t0);
    if ((rc = c.ensure(&c.scratch_b, &c.scratch_b_bytes, 2 * bbytes))) return finish(info, rc, t0);
    if ((rc = c.ensure(&c.scratch_perm, &c.scratch_perm_bytes, (size_t)n * sizeof(int64_t)))) return finish(info, rc, t0);
    T *dA = (T *)c.scratch_a, *dB = (T *)c.scratch_b, *dX = dB + (size_t)n * nrhs;
    if ((rc = scirs::copy_h2d(dA, n * sizeof(T), a, lda * sizeof(T), n * sizeof(T), n, s))) return finish(info, rc, t0);
    if ((rc = scirs::copy_h2d(dB, nrhs * sizeof(T), b, ldb * sizeof(T), nrhs * sizeof(T), n, s)))
        return finish(info, rc, t0);
    int st = getrf_solve_device<T>(n, nrhs, dA, n, (int64_t *)c.scratch_perm, dB, nrhs, dX, nrhs, s, info);
    if (st < 0) return finish(info, st, t0);
    if (st == 0) {
        if ((rc = scirs::copy_d2h(x, ldx * sizeof(T), dX, nrhs * sizeof(T), nrhs * sizeof(T), n, s)))
            return finish(info, rc, t0);
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    }
    return finish(info, st, t0);
}

extern "C" int scirs_gpu_dgesv(int n, int nrhs, const double *a, int64_t lda, const double *b, int64_t ldb, double *x,
                               int64_t ldx, scirs_solve_info *info)
{
    return host_gesv<double>(n, nrhs, a, lda, b, ldb, x, ldx, info);
}

extern "C" int scirs_gpu_sgesv(int n, int nrhs, const float *a, int64_t lda, const float *b, int64_t ldb, float *x,
                               int64_t ldx, scirs_solve_info *info)
{
    return host_gesv<float>(n, nrhs, a, lda, b, ldb, x, ldx, info);
}

extern "C" int scirs_gpu_dgesv_dev(int n, int nrhs, const double *dA, int64_t lda, const double *dB, int64_t ldb,
                                   double *dX, int64_t ldx, double *dLU, int64_t ldlu, int64_t *dPerm, void *stream,
                                   scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    if (n <= 0 || nrhs < 0 || !dA || lda < n || (nrhs > 0 && (!dB || !dX || ldb < nrhs || ldx < nrhs))) {
        scirs_set_error("dgesv_dev: invalid arguments (n=%d nrhs=%d)", n, nrhs);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    SCIRS_ENTER(true);
    Context &c = ctx();
    hipStream_t s = (hipStream_t)stream;   // NULL: the default stream, ordered after the caller's work on it
    int rc;
    if (!dLU) {
        if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, (size_t)n * n * sizeof(double)))) return finish(info, rc, t0);
        dLU = (double *)c.scratch_a;
        ldlu = n;
    }
    if (!dPerm) {
        if ((rc = c.ensure(&c.scratch_perm, &c.scratch_perm_bytes, (size_t)n * sizeof(int64_t)))) return finish(info, rc, t0);
        dPerm = (int64_t *)c.scratch_perm;
    }
    // the reference factors a copy of A (lapack.rs:86); so does this entry point
    if ((rc = scirs_d2d_copy(n, n, dA, lda, dLU, ldlu, s))) return finish(info, rc, t0);
    int st = getrf_solve_device<double>(n, nrhs, dLU, ldlu, dPerm, dB, ldb, dX, ldx, s, info);
    return finish(info, st, t0);
}

// ------------------------------------------------------------------ Cholesky
template <typename T>
static int host_potrf(int n, T *a, int64_t lda, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, sizeof(T) == 8 ? SCIRS_PATH_FP64_CHOL : SCIRS_PATH_FP32_CHOL);
    if (n <= 0 || !a || lda < n) {
        scirs_set_error("potrf: invalid arguments (n=%d)", n);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    SCIRS_ENTER(true);
    Context &c = ctx();
    hipStream_t s = c.stream;
    int rc;
    if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, (size_t)n * n * sizeof(T)))) return finish(info, rc, t0);
    T *dA = (T *)c.scratch_a;
    if ((rc = scirs::copy_h2d(dA, n * sizeof(T), a, lda * sizeof(T), n * sizeof(T), n, s))) return finish(info, rc, t0);
    if ((rc = c.reset_flags(s))) return finish(info, rc, t0);
    if ((rc = scirs_potrf_device<T>(n, dA, n, c.kb_chol, c.work, s))) return finish(info, rc, t0);
    int st = 0;
    double pv = 0;
    if ((rc = c.read_flags(s, &st, &pv))) return finish(info, rc, t0);
    if (info) info->pivot_value = st ? pv : 0.0;
    if (st == 0) {
        if ((rc = scirs::copy_d2h(a, lda * sizeof(T), dA, n * sizeof(T), n * sizeof(T), n, s)))
            return finish(info, rc, t0);
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    }
    return finish(info, st, t0);
}

extern "C" int scirs_gpu_dpotrf(int n, double *a, int64_t lda, scirs_solve_info *info)
{
    return host_potrf<double>(n, a, lda, info);
}

extern "C" int scirs_gpu_spotrf(int n, float *a, int64_t lda, scirs_solve_info *info)
{
    return host_potrf<float>(n, a, lda, info);
}

template <typename T>
static int posv_device(int n, int nrhs, const T *dA, long lda, const T *dB, long ldb, T *dX, long ldx, T *dL,
                       long ldl, hipStream_t s, scirs_solve_info *info)
{
    Context &c = ctx();
    int rc;
    scirs::CallEvents ev;
    if (int erc = ev.create(3)) return erc;
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1], e2 = ev.e[2];
    SCIRS_HIP_CHECK(hipEventRecord(e0, s));
    if ((rc = scirs_d2d_copy(n, n, dA, lda, dL, ldl, s))) return rc;
    if ((rc = c.reset_flags(s))) return rc;
    if ((rc = scirs_potrf_device<T>(n, dL, ldl, c.kb_chol, c.work, s))) return rc;
    SCIRS_HIP_CHECK(hipEventRecord(e1, s));
    if ((rc = scirs_d2d_copy(n, nrhs, dB, ldb, dX, ldx, s))) return rc;
    if (c.chol_chain) {
        // L y = b and L^T x = y as chained solves (true division by the diagonal,
        // solve.rs:159-199), the second on an explicit row-major L^T
        if ((rc = c.ensure(&c.scratch_t, &c.scratch_t_bytes, (size_t)n * n * sizeof(T)))) return rc;
        T *dU = (T *)c.scratch_t;
        if ((rc = scirs_transpose_lower<T>(n, dL, ldl, dU, n, s))) return rc;
        if ((rc = c.ensure_xtri(n))) return rc;
        if ((rc = scirs_trsv_chain<T>(n, nrhs, dL, ldl, dX, ldx, 1, 0, c.work, s))) return rc;
        if ((rc = scirs_trsv_chain<T>(n, nrhs, dU, n, dX, ldx, 0, 0, c.work, s))) return rc;
    } else if ((rc = scirs_potrs_device<T>(n, nrhs, dL, ldl, dX, ldx, c.work.info, s))) {
        return rc;
    }
    SCIRS_HIP_CHECK(hipEventRecord(e2, s));
    int st = 0;
    double pv = 0;
    if ((rc = c.read_flags(s, &st, &pv))) return rc;
    if (info) {
        float ms1 = 0, ms2 = 0;
        (void)hipEventElapsedTime(&ms1, e0, e1);
        (void)hipEventElapsedTime(&ms2, e1, e2);
        info->time_factor_ms = ms1;
        info->time_solve_ms = ms2;
        info->pivot_value = st ? pv : 0.0;
    }
    return st;
}

template <typename T>
static int host_posv(int n, int nrhs, const T *a, int64_t lda, const T *b, int64_t ldb, T *x, int64_t ldx,
                     scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_CHOL);
    if (n <= 0 || nrhs <= 0 || !a || !b || !x || lda < n || ldb < nrhs || ldx < nrhs) {
        scirs_set_error("%cposv: invalid arguments (n=%d nrhs=%d)", sizeof(T) == 8 ? 'd' : 's', n, nrhs);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    SCIRS_ENTER(true);
    Context &c = ctx();
    hipStream_t s = c.stream;
    int rc;
    const size_t abytes = (size_t)n * n * sizeof(T), bbytes = (size_t)n * nrhs * sizeof(T);
    if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, 2 * abytes))) return finish(info, rc, t0);
    if ((rc = c.ensure(&c.scratch_b, &c.scratch_b_bytes, 2 * bbytes))) return finish(info, rc, t0);
    T *dA = (T *)c.scratch_a, *dL = dA + (size_t)n * n;
    T *dB = (T *)c.scratch_b, *dX = dB + (size_t)n * nrhs;
    if ((rc = scirs::copy_h2d(dA, n * sizeof(T), a, lda * sizeof(T), n * sizeof(T), n, s))) return finish(info, rc, t0);
    if ((rc = scirs::copy_h2d(dB, nrhs * sizeof(T), b, ldb * sizeof(T), nrhs * sizeof(T), n, s)))
        return finish(info, rc, t0);
    int st = posv_device<T>(n, nrhs, dA, n, dB, nrhs, dX, nrhs, dL, n, s, info);
    if (st < 0) return finish(info, st, t0);
    if (st == 0) {
        if ((rc = scirs::copy_d2h(x, ldx * sizeof(T), dX, nrhs * sizeof(T), nrhs * sizeof(T), n, s)))
            return finish(info, rc, t0);
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    }
    return finish(info, st, t0);
}

extern "C" int scirs_gpu_dposv(int n, int nrhs, const double *a, int64_t lda, const double *b, int64_t ldb, double *x,
                               int64_t ldx, scirs_solve_info *info)
{
    return host_posv<double>(n, nrhs, a, lda, b, ldb, x, ldx, info);
}

extern "C" int scirs_gpu_sposv(int n, int nrhs, const float *a, int64_t lda, const float *b, int64_t ldb, float *x,
                               int64_t ldx, scirs_solve_info *info)
{
    return host_posv<float>(n, nrhs, a, lda, b, ldb, x, ldx, info);
}

extern "C" int scirs_gpu_dposv_dev(int n, int nrhs, const double *dA, int64_t lda, const double *dB, int64_t ldb,
                                   double *dX, int64_t ldx, double *dL, int64_t ldl, void *stream,
                                   scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_CHOL);
    if (n <= 0 || nrhs <= 0 || !dA || !dB || !dX || lda < n || ldb < nrhs || ldx < nrhs) {
        scirs_set_error("dposv_dev: invalid arguments (n=%d nrhs=%d)", n, nrhs);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    SCIRS_ENTER(true);
    Context &c = ctx();
    hipStream_t s = (hipStream_t)stream;   // NULL: the default stream, ordered after the caller's work on it
    int rc;
    if (!dL) {
        if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, (size_t)n * n * sizeof(double)))) return finish(info, rc, t0);
        dL = (double *)c.scratch_a;
        ldl = n;
    }
    int st = posv_device<double>(n, nrhs, dA, lda, dB, ldb, dX, ldx, dL, ldl, s, info);
    return finish(info, st, t0);
}

// ------------------------------------------------------------------ triangular
template <typename T>
static int host_trsv(int n, int nrhs, const T *a, int64_t lda, const T *b, int64_t ldb, T *x, int64_t ldx, int lower,
                     int unit_diagonal, scirs_solve_info *info)
{
    const double t0 = now_ms();
    info_init(info, SCIRS_PATH_FP64_LU);
    if (n <= 0 || nrhs <= 0 || !a || !b || !x || lda < n || ldb < nrhs || ldx < nrhs) {
        scirs_set_error("%ctrsv: invalid arguments (n=%d nrhs=%d)", sizeof(T) == 8 ? 'd' : 's', n, nrhs);
        return finish(info, SCIRS_ERR_ARG, t0);
    }
    // |a_ii| < eps (of T) on a non-unit diagonal is singular (solve.rs:163-170, 184-191);
    // the first such row in substitution order is reported.
    if (!unit_diagonal) {
        for (int k = 0; k < n; ++k) {
            int i = lower ? k : n - 1 - k;
            const double d = (double)a[(size_t)i * lda + i];
            if (fabs(d) < (sizeof(T) == 8 ? DBL_EPSILON : (double)FLT_EPSILON)) {
                if (info) info->pivot_value = fabs(d);
                return finish(info, i + 1, t0);
            }
        }
    }
    SCIRS_ENTER(true);
    Context &c = ctx();
    hipStream_t s = c.stream;
    int rc;
    if ((rc = c.ensure(&c.scratch_a, &c.scratch_a_bytes, (size_t)n * n * sizeof(T)))) return finish(info, rc, t0);
    if ((rc = c.ensure(&c.scratch_b, &c.scratch_b_bytes, (size_t)n * nrhs * sizeof(T)))) return finish(info, rc, t0);
    T *dA = (T *)c.scratch_a, *dX = (T *)c.scratch_b;
    if ((rc = scirs::copy_h2d(dA, n * sizeof(T), a, lda * sizeof(T), n * sizeof(T), n, s))) return finish(info, rc, t0);
    if ((rc = scirs::copy_h2d(dX, nrhs * sizeof(T), b, ldb * sizeof(T), nrhs * sizeof(T), n, s)))
        return finish(info, rc, t0);
    if ((rc = c.ensure_xtri(n))) return finish(info, rc, t0);
    if ((rc = c.reset_flags(s))) return finish(info, rc, t0);
    if ((rc = scirs_trsv_chain<T>(n, nrhs, dA, n, dX, nrhs, lower, unit_diagonal, c.work, s)))
        return finish(info, rc, t0);
    int st = 0;
    if ((rc = c.read_flags(s, &st, nullptr))) return finish(info, rc, t0);
    if ((rc = scirs::copy_d2h(x, ldx * sizeof(T), dX, nrhs * sizeof(T), nrhs * sizeof(T), n, s)))
        return finish(info, rc, t0);
    SCIRS_HIP_CHECK(hipStreamSynchronize(s));
    return finish(info, SCIRS_OK, t0);
}

extern "C" int scirs_gpu_dtrsv(int n, int nrhs, const double *a, int64_t lda, const double *b, int64_t ldb, double *x,
                               int64_t ldx, int lower, int unit_diagonal, scirs_solve_info *info)
{
    return host_trsv<double>(n, nrhs, a, lda, b, ldb, x, ldx, lower, unit_diagonal, info);
}

extern "C" int scirs_gpu_strsv(int n, int nrhs, const float *a, int64_t lda, const float *b, int64_t ldb, float *x,
                               int64_t ldx, int lower, int unit_diagonal, scirs_solve_info *info)
{
    return host_trsv<float>(n, nrhs, a, lda, b, ldb, x, ldx, lower, unit_diagonal, info);
}

// ------------------------------------------------------------------ diagnostics
extern "C" int scirs_gpu_backward_error_dev(int n, int nrhs, const double *dA, int64_t lda, const double *dX,
                                            int64_t ldx, const double *dB, int64_t ldb, void *stream, double *out)
{
    if (n <= 0 || nrhs <= 0 || !dA || !dX || !dB || !out) return SCIRS_ERR_ARG;
    SCIRS_ENTER(true);
    hipStream_t s = (hipStream_t)stream;   // NULL: the default stream, ordered after the caller's work on it
    return backward_error(n, nrhs, dA, lda, dX, ldx, dB, ldb, s, out);
}

extern "C" int scirs_gpu_backward_error_dev_noalloc(int n, int nrhs, const double *dA, int64_t lda,
                                                    const double *dX, int64_t ldx, const double *dB, int64_t ldb,
                                                    void *stream, double *out)
{
    if (n <= 0 || nrhs <= 0 || !dA || !dX || !dB || !out || lda < n || ldx < nrhs || ldb < nrhs) {
        scirs_set_error("backward_error_dev_noalloc: invalid arguments (n=%d nrhs=%d)", n, nrhs);
        return SCIRS_ERR_ARG;
    }
    if (backward_error_scratch(n, nrhs) > SCIRS_NOALLOC_SCRATCH_DOUBLES) {
        scirs_set_error("backward_error_dev_noalloc: n * nrhs + n + 200 = %zu doubles exceed its fixed scratch of %u "
                        "(use scirs_gpu_backward_error_dev)", backward_error_scratch(n, nrhs),
                        (unsigned)SCIRS_NOALLOC_SCRATCH_DOUBLES);
        return SCIRS_ERR_ARG;
    }
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    Context &c = ctx();
    // initialising the library allocates, so it must have happened before
    if (!c.initialised || !c.noalloc_scratch) {
        scirs_set_error("backward_error_dev_noalloc: call scirs_gpu_init first (it sets up the fixed scratch)");
        return SCIRS_ERR_UNSUPPORTED;
    }
    SCIRS_HIP_CHECK(hipSetDevice(c.device));
    return backward_error_in(n, nrhs, dA, lda, dX, ldx, dB, ldb, (hipStream_t)stream, c.noalloc_scratch, out);
}

// Clear every sequence-tagged exchange word (a kernel, so it is safe under capture).
static int clear_tags(ScirsWork &w, hipStream_t s)
{
    struct Buf { unsigned long long *p; size_t words; };
    const Buf bufs[] = {{w.xw, w.xw_words}, {w.xrow, w.xrow_words}, {w.xrj, w.xrj_words}, {w.xtri, w.xtri_words},
                        {w.tp_cand, (size_t)TSLU_MAXLEV * TSLU_MAXNODE * 65}};
    for (const Buf &b : bufs) {
        if (!b.p || !b.words) continue;
        for (size_t off = 0; off < 2 * b.words; off += (size_t)1 << 30) {
            const size_t cnt = std::min((size_t)1 << 30, 2 * b.words - off);
            if (int rc = scirs_fill_int((int *)b.p + off, (int)cnt, 0, s)) return rc;
        }
    }
    return SCIRS_OK;
}

int scirs_next_seq(ScirsWork &w, unsigned span, hipStream_t s, unsigned *seq)
{
    unsigned next = w.seq_next, base = 0;
    if (scirs_gpu_seq_plan(next, span, &base, &next)) {
        if (int rc = clear_tags(w, s)) return rc;
    }
    w.seq_next = next;
    *seq = base;
    return SCIRS_OK;
}

extern "C" int scirs_gpu_seq_plan(unsigned next, unsigned span, unsigned *base, unsigned *next_out)
{
    // a launch tags its words with base .. base + span - 1 (never 0, the value of
    // a cleared word); when that range would reach the top of the 32-bit count,
    // the caller clears the tagged buffers and the count restarts at 1
    int reset = 0;
    if (next == 0u || span == 0u || next > 0xFFFFFFFFu - span) {
        next = 1u;
        reset = 1;
    }
    *base = next;
    *next_out = next + (span ? span : 1u);
    return reset;
}
