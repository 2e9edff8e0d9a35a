// Process-wide context and helpers shared by the C-ABI translation units.
#pragma once
#include "common.hpp"
#include "kernels.hpp"
#include "../../include/scirs_gpu.h"
#include <mutex>
#include <algorithm>
#include <cmath>

namespace scirs {

struct Context {
    std::recursive_mutex mu;
    bool initialised = false;
    int device = -1;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    ScirsWork work{};
    void *work_base = nullptr;
    void *host_flags = nullptr;
    int kb_f64 = 128, kb_f32 = 128, kb_chol = 128;
    int narrow_kb = 1;   // narrower LU panels when a whole panel then fits one workgroup
    int chol_chain = 1;  // Cholesky solves through the chained triangular solve (else blocked)
    int trsv_dinv = 1;   // FP32 refinement solves with inverted 64 x 64 diagonal blocks
    // LU panel width for an m-row factorisation in element type T
    template <typename T> int lu_kb(int m) const
    {
        const int kb = sizeof(T) == 8 ? kb_f64 : kb_f32;
        return narrow_kb ? scirs_lu_panel_kb(m, kb, sizeof(T)) : kb;
    }
    // scratch grown on demand (host-memory entry points and internal temporaries)
    void *scratch_a = nullptr; size_t scratch_a_bytes = 0;
    void *scratch_b = nullptr; size_t scratch_b_bytes = 0;
    void *scratch_perm = nullptr; size_t scratch_perm_bytes = 0;
    void *scratch_bwd = nullptr; size_t scratch_bwd_bytes = 0;
    void *scratch_ir = nullptr; size_t scratch_ir_bytes = 0;
    void *scratch_ir2 = nullptr; size_t scratch_ir2_bytes = 0;
    void *scratch_t = nullptr; size_t scratch_t_bytes = 0;   // L^T of a Cholesky solve
    void *scratch_lu64 = nullptr; size_t scratch_lu64_bytes = 0;   // FP64 copy of GMRES-IR's low-precision factors
    void *xtri_mem = nullptr; size_t xtri_bytes = 0;
    void *colexp_mem = nullptr; size_t colexp_bytes = 0;
    double *noalloc_scratch = nullptr;   // device address of the fixed scratch array (set at init)

    int init(int dev);
    int ensure(void **p, size_t *have, size_t bytes);
    int reset_flags(hipStream_t s);
    int read_flags(hipStream_t s, int *info, double *info_val);
    // exchange buffer of the chained triangular solve for systems of order n
    int ensure_xtri(int n);
};

// X <- U^-1 L^-1 X with the packed LU factors (X already row-permuted);
// dinv_l / dinv_u (optional): inverses of the 64 x 64 diagonal blocks of L and U;
// dmt_l / dmt_u (optional, with them): those inverses times each block row's last tile.
template <typename T>
int lu_substitute(int n, int nrhs, const T *dLU, long ldlu, T *dX, long ldx, hipStream_t s,
                  const T *dinv_l = nullptr, const T *dinv_u = nullptr, const T *dmt_l = nullptr,
                  const T *dmt_u = nullptr);

// Timing events of one call, destroyed on every return path.
struct CallEvents {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    int create(int n)
    {
        for (int i = 0; i < n; ++i)
            if (hipEventCreate(&e[i]) != hipSuccess) {
                scirs_set_error("hipEventCreate failed");
                return SCIRS_ERR_HIP;
            }
        return SCIRS_OK;
    }
    ~CallEvents()
    {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

Context &ctx();
double now_ms();
void info_init(scirs_solve_info *info, int path);
int finish(scirs_solve_info *info, int rc, double t0);
int backward_error(int n, int nrhs, const double *dA, long lda, const double *dX, long ldx, const double *dB,
                   long ldb, hipStream_t s, double *out);

// hostcopy.hip: rows x width bytes between a strided host array (pitch hpitch)
// and device memory (pitch dpitch) on stream s, through pinned staging.
// copy_h2d returns once the host array has been read (the last transfer may
// still run on s); copy_d2h returns once the host array has been written.
int copy_h2d(void *d, size_t dpitch, const void *h, size_t hpitch, size_t width, size_t rows, hipStream_t s);
int copy_d2h(void *h, size_t hpitch, const void *d, size_t dpitch, size_t width, size_t rows, hipStream_t s);
// the same for one contiguous block of bytes
int copy_h2d_flat(void *d, const void *h, size_t bytes, hipStream_t s);
int copy_d2h_flat(void *h, const void *d, size_t bytes, hipStream_t s);

} // namespace scirs

// vec.hip
int scirs_residual(int n, int nrhs, const double *A, long lda, const double *X, long ldx, const double *B, long ldb,
                   double *R, long ldr, hipStream_t s);
int scirs_colmax_d(int n, int ncols, const double *V, long ldv, double *out, hipStream_t s);
int scirs_colmax_s(int n, int ncols, const float *V, long ldv, double *out, hipStream_t s);
int scirs_convert_rowsum(int n, int m, const double *A, long lda, float *A32, long lda32, double *rowsum,
                         hipStream_t s);
int scirs_rowsum_max(int n, const double *rowsum, double *out, hipStream_t s);
int scirs_d2s(int rows, int cols, const double *src, long lds, float *dst, long ldd, hipStream_t s);
int scirs_s2d(int rows, int cols, const float *src, long lds, double *dst, long ldd, hipStream_t s);
int scirs_d2d_copy(int rows, int cols, const double *src, long lds, double *dst, long ldd, hipStream_t s);
int scirs_d2d_copy(int rows, int cols, const float *src, long lds, float *dst, long ldd, hipStream_t s);
int scirs_add_s_to_d(int rows, int cols, const float *D, long ldd, double *X, long ldx, hipStream_t s);
int scirs_add_d_to_d(int rows, int cols, const double *D, long ldd, double *X, long ldx, hipStream_t s);
int scirs_fill_int(int *p, int count, int value, hipStream_t s);
