// Internal launchers shared between the kernel files and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// C[M x N] -= A[M x K] * op(B); op(B) = B (K x N) or, when trans_b, B stored N x K.
// lower != 0 skips 128x128 tiles strictly above C's diagonal.  Does nothing when *info != 0.
int scirs_launch_dgemm_sub(int M, int N, int K, const double *A, long lda, const double *B, long ldb,
                           int trans_b, double *C, long ldc, int lower, const int *info,
                           hipStream_t stream);
// C = A B^T for the Cholesky TRSM: C[M x N] = A[M x 128] B^T with B N x 128,
// N <= 128; C may alias A (each workgroup owns whole rows).
int scirs_launch_dgemm_assign_t(int M, int N, int K, const double *A, long lda, const double *B, long ldb,
                                double *C, long ldc, const int *info, hipStream_t stream);

// Device scratch used by the factorisations (allocated once per context).
struct ScirsWork {
    unsigned long long *xw;    // [REP][2][MAX_G][4]        panel candidates (sequence-tagged words)
    unsigned long long *xrow;  // [REP][2][MAX_G][KB_MAX*2] panel candidate rows
    unsigned long long *xrj;   // [REP][2][KB_MAX*2]        panel row j
    int panel_nrep;            // replicas of the panel's published words (1..SCIRS_PANEL_REP_MAX)
    unsigned seq_next;         // sequence base of the next panel launch (host side)
    size_t xw_words, xrow_words, xrj_words;
    int *abort;          // raised by a timed-out exchange
    int *pairs;          // [n_panels_max][2*KB_MAX][2] row gather lists (dst, src), absolute
    int *npairs;         // [n_panels_max]
    int *info;           // 1-based failing column / row, 0 = ok
    double *info_val;    // |pivot| (LU) or a_ii - sum (Cholesky) at the failure
    int max_panels;
    int max_g;           // co-resident workgroup limit for cooperative launches
    unsigned spin_limit;
    int lowp_mode;       // trailing-update arithmetic of the FP32 factorisation (SCIRS_LOWP_*)
    int panel_rows;      // target rows per workgroup in the cooperative panel
    unsigned long long *xtri;  // chained triangular solve: sequence-tagged x values
    size_t xtri_words;
    int *ticket;         // chained triangular solve: block-row ticket counter
    void *linv;          // 128 x 128 inverse of a panel's unit-lower block (FP32 TRSM)
    double *linv64;      // 128 x 128 inverse of the Cholesky diagonal block (FP64 TRSM as a product)
    int panel_w0;        // panel critical path on wave 0 alone (1) or on all waves (0)
    int *colexp;         // FP16 update: per-column exponents of U12, two regions (main / side stream)
    int colexp_half;     // columns per region
    int lookahead;       // overlap panel k+1 with the rest of update k on side_stream
    int trsv_pair;       // FP32 one-rhs chained solves with couplings: two block rows per link
    int poll_sleep;      // back-off of the panel's candidate poll (s_sleep(2) repetitions)
    hipStream_t side_stream;
    hipEvent_t ev_trsm, ev_upd, ev_start;
    // tournament-pivoted panel (FP32 factorisation of the mixed solve)
    int tslu;                  // the mixed solve's FP32 factorisation pivots by tournament (1) or exactly (0)
    int tslu_active;           // set by the mixed solve around its factorisation (square, lda % 4 == 0)
    unsigned long long *tp_cand;   // [TSLU_MAXLEV][TSLU_MAXNODE][64 + 1] candidate lists (sequence-tagged words)
    int *tp_cnt;               // [TSLU_MAXLEV][TSLU_MAXNODE] arrival counters (zero between launches)
    float *tp_uinv;            // 64 x 64 inverse of a sub-panel's U block
    float *tp_urest;           // 64 x 64 the sub-panel's U rows on the panel's later columns
    float *tp_top;             // 64 x SCIRS_KB_MAX final contents of the sub-panel's pivot rows
    float *tp_stage;           // 64 x SCIRS_KB_MAX rows displaced from the sub-panel's top positions
    int *tp_vlist;             // [64 + 1] positions the displaced rows move to, then their count
    int *tp_map;               // [2 * TSLU_MAP] (position, original row) of the panel's moved rows, then a count
};

enum { TSLU_MAXLEV = 6, TSLU_MAXNODE = 512, TSLU_MAP = 512 };

// Sequence number base for a launch that uses `span` consecutive tags of the
// exchange buffers.  Before the 32-bit counter would wrap, every tagged buffer
// is cleared on `s` and the count restarts at 1, so a stale word can never
// carry a current tag.
int scirs_next_seq(ScirsWork &w, unsigned span, hipStream_t s, unsigned *seq);

// The tournament-pivoted factorisation of the kb-column panel at k0 of the n x n
// FP32 matrix A (lda a multiple of 4): the panel's rows are permuted, the panel is
// factored in place (unit L below, U on and above the diagonal) and w.pairs / w.npairs
// receive its row interchanges, as the partial-pivot panel leaves them.
int scirs_tslu_panel(float *A, long lda, int n, int k0, int kb, int panel_index, ScirsWork &w, hipStream_t s);

enum {
    SCIRS_LOWP_FP32 = 0,   // FP32 operands (VALU FMA)
    SCIRS_LOWP_BF16 = 1,   // BF16 operands, FP32 accumulation (MFMA)
    SCIRS_LOWP_FP16 = 2,   // FP16 operands with per-block power-of-two scaling, FP32 accumulation
    SCIRS_LOWP_BF16X3 = 3, // FP32 split into hi+lo BF16 pairs, three MFMA products
    SCIRS_LOWP_BF16X6 = 4, // FP32 split into hi+mid+lo BF16 terms, six MFMA products (FP32-accurate)
    SCIRS_LOWP_FP16X3 = 5, // FP32 split into hi+lo FP16 pairs (power-of-two scaled), three MFMA products
};

enum { SCIRS_KB_MAX = 256, SCIRS_MAX_G = 512, SCIRS_PANEL_REP_MAX = 8 };
// chains of a chained triangular solve launched together (one ticket each, in the flags block)
enum { SCIRS_CHAIN_GROUPS = 32 };
// value of *abort raised by a timed-out QR panel exchange (the LU panel raises
// its column + 1, a chained triangular solve -1)
enum { SCIRS_ABORT_QR_PANEL = -2 };

// Trailing (Schur) update C -= A B of the factorisations: FP64 MFMA for double,
// the arithmetic chosen by w.lowp_mode for float.
template <typename T>
int scirs_trailing_update(int M, int N, int K, const T *A, long lda, const T *B, long ldb, T *C, long ldc,
                          const ScirsWork &w, hipStream_t stream);
// C (N x N, lower tiles) -= A A^T with A N x K.
template <typename T>
int scirs_syrk_update(int N, int K, const T *A, long lda, T *C, long ldc, const ScirsWork &w,
                      hipStream_t stream);
// C (M x N) -= A (M x K) B^T with B N x K (the Cholesky trailing update below a corner)
template <typename T>
int scirs_launch_gemm_sub_t(int M, int N, int K, const T *A, long lda, const T *B, long ldb, T *C, long ldc,
                            const ScirsWork &w, hipStream_t stream);

template <typename T>
int scirs_potrf_device(int n, T *A, long lda, int kb, ScirsWork &w, hipStream_t stream);
template <typename T>
int scirs_potrs_device(int n, int nrhs, const T *L, long ldl, T *B, long ldb, const int *info,
                       hipStream_t stream);
template <typename T>
int scirs_trsv_device(int n, int nrhs, const T *A, long lda, T *B, long ldb, int lower, int unit,
                      const int *info, hipStream_t stream);

// FP32 update C -= A * op(B) in the arithmetic of `mode`.
int scirs_launch_sgemm_sub(int M, int N, int K, const float *A, long lda, const float *B, long ldb,
                           int trans_b, float *C, long ldc, int lower, int mode, const int *info,
                           hipStream_t stream);

// FP16 update with power-of-two scalings: A times 2^a_shift, column c of B times
// 2^-colexp[c] (both undone exactly in FP32); colexp from scirs_launch_colexp.
int scirs_launch_sgemm_fp16_scaled(int M, int N, int K, const float *A, long lda, const float *B, long ldb,
                                   float *C, long ldc, const int *colexp, int a_shift, int mode, const int *info,
                                   hipStream_t stream);
int scirs_launch_colexp(int K, int N, const float *B, long ldb, int *colexp, const int *info, hipStream_t stream);

// Triangular solve with 256-row diagonal blocks (two launches per block):
// X <- op(A)^-1 X with op the lower (unit or not) or upper triangle of A.
template <typename T>
int scirs_trsv_big(int n, int nrhs, const T *A, long lda, T *X, long ldx, int lower, int unit, const int *info,
                   hipStream_t s);

// Chained triangular solve in one launch (64-row block rows handed over through
// w.xtri, which must hold n * 4 * 2 words): X <- op(A)^-1 X as scirs_trsv_big.
// dinv (optional): inverses of the 64 x 64 diagonal blocks (scirs_tri_block_inverse),
// applied instead of the substitution inside each block
// dmt (optional, with dinv): Dinv[p] times the tile block row p meets last
// (scirs_tri_block_coupling, depth 3: a one-rhs FP32 solve then hands over
// two block rows per link, using the couplings two and three blocks back)
template <typename T>
int scirs_trsv_chain(int n, int nrhs, const T *A, long lda, T *X, long ldx, int lower, int unit, ScirsWork &w,
                     hipStream_t s, const T *dinv = nullptr, const T *dmt = nullptr);

// M_d[p] = Dinv[p] * (block row p's d-th from last tile in the chained solve's
// order) for d = 1 .. depth (<= 3): depth * ceil(n/64) * 64 * 64 floats, M_d at
// (d - 1) * ceil(n/64) * 64 * 64
int scirs_tri_block_coupling(int n, const float *A, long lda, int lower, const float *dinv, float *dmt, int depth,
                             hipStream_t s);

// Dinv[p] = (T_pp)^-1 for the 64 x 64 diagonal blocks of the lower (unit or
// not) or upper triangle of the FP32 matrix A; dinv holds ceil(n/64) * 64 * 64 floats
int scirs_tri_block_inverse(int n, const float *A, long lda, int lower, int unit, float *dinv, hipStream_t s);

template <typename T>
int scirs_getrf_device(int m, int n, T *A, long lda, int kb, ScirsWork &w, hipStream_t stream);

// U = L^T (upper triangle of U written; L's strict upper triangle must be zero)
template <typename T>
int scirs_transpose_lower(int n, const T *L, long ldl, T *U, long ldu, hipStream_t s);

// Panel width for an m-row factorisation: the widest of kb, 64, 32 (never above
// kb) at which all m rows of a panel fit one workgroup's LDS, so every panel is
// factored by one workgroup without inter-workgroup exchange; kb when none fits.
int scirs_lu_panel_kb(int m, int kb, size_t elem_bytes);

// Row permutation produced by the last getrf: perm[i] = original row now at i
// (the `piv` vector of scirs2-linalg's LUDecomposition).
int scirs_perm_from_pairs(int m, int n, int kb, const ScirsWork &w, int64_t *perm, hipStream_t stream);

// X[i, :] = B[perm[i], :] for an n x nrhs row-major block.
template <typename T>
int scirs_permute_rows(int n, int nrhs, const int64_t *perm, const T *B, long ldb, T *X, long ldx,
                       hipStream_t stream);

// Solve L U X = B in place on B (B already row-permuted); L unit lower, U upper, packed in A.
template <typename T>
int scirs_getrs_device(int n, int nrhs, const T *A, long lda, T *B, long ldb, const int *info,
                       hipStream_t stream);
