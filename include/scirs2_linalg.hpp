// scirs2_linalg.hpp — C++17 mirror of the scirs2-linalg dense-solve API on top of
// the C ABI in scirs_gpu.h.  It is the host side a binding keeps: the same
// function names, argument meaning, validation order, error variants and error
// texts as the Rust functions it stands in for, with the O(n^3) work sent to the
// GPU library.
//
//   solve            scirs2-linalg/src/solve.rs:54-111
//   solve_triangular scirs2-linalg/src/solve.rs:139-202
//   solve_multiple   scirs2-linalg/src/solve.rs:366-424
//   lu               scirs2-linalg/src/decomposition.rs:108-183
//   cholesky         scirs2-linalg/src/decomposition.rs:44-81
//   inv, det         scirs2-linalg/src/basic.rs:117-186, 29-93
//   mixed_precision_solve, iterative_refinement_solve
//                    scirs2-linalg/src/mixed_precision/adaptive.rs:49-151, 282-378
//   compat_solve     scirs2-linalg/src/compat.rs:461-499
//   qr, lstsq        scirs2-linalg/src/decomposition.rs:208-249, solve.rs:234-290 (rows >= cols)
//   gmres            scirs2-linalg/src/solvers/iterative.rs:317-450
//   preconditioned_gmres, PreconditionerOp, DiagonalPreconditioner
//                    scirs2-linalg/src/preconditioners.rs:179-245, 934-1010
//   batch_matmul, batch_matvec, batch_add, batch_sum
//                    scirs2-linalg/src/batch/mod.rs:62-336
//   validation       scirs2-linalg/src/validation.rs:22-329
//   LinalgError      scirs2-linalg/src/error.rs:8-150
//
// Header-only; link with libscirs_gpu.so.  Element types: double and float.
#ifndef SCIRS2_LINALG_HPP
#define SCIRS2_LINALG_HPP

#include "scirs_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <optional>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <type_traits>
#include <utility>
#include <vector>

namespace scirs2_linalg {

// ------------------------------------------------------------------ errors
// Variants and Display prefixes of LinalgError (error.rs:8-66).
class LinalgError {
  public:
    enum class Kind {
        ComputationError, DomainError, ConvergenceError, DimensionError, ShapeError, IndexError,
        SingularMatrixError, NonPositiveDefiniteError, NotImplementedError, ImplementationError,
        ValueError, InvalidInputError, InvalidInput, NumericalError
    };
    LinalgError(Kind k, std::string msg) : kind_(k), msg_(std::move(msg)) {}
    Kind kind() const { return kind_; }
    const std::string &message() const { return msg_; }
    // the Display text of the Rust enum
    std::string to_string() const
    {
        static const char *prefix[] = {
            "Computation error: ", "Domain error: ", "Convergence error: ", "Dimension mismatch error: ",
            "Shape error: ", "Index out of bounds: ", "Singular matrix error: ",
            "Non-positive definite matrix error: ", "Not implemented: ", "Implementation error: ",
            "Value error: ", "Invalid input error: ", "Invalid input: ", "Numerical error: "};
        return std::string(prefix[(int)kind_]) + msg_;
    }

  private:
    Kind kind_;
    std::string msg_;
};

inline std::string shape_str(std::size_t r, std::size_t c)
{
    return std::to_string(r) + "\xC3\x97" + std::to_string(c);   // "r×c"
}

// {:.2e} of Rust for the condition-number line
inline std::string fmt_e2(double v)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.2e", v);
    // Rust prints the exponent without sign padding: 1.00e16, 2.50e-3
    std::string s(buf);
    auto e = s.find('e');
    if (e == std::string::npos) return s;
    std::string mant = s.substr(0, e), ex = s.substr(e + 1);
    bool neg = !ex.empty() && ex[0] == '-';
    if (!ex.empty() && (ex[0] == '+' || ex[0] == '-')) ex = ex.substr(1);
    while (ex.size() > 1 && ex[0] == '0') ex = ex.substr(1);
    return mant + "e" + (neg ? "-" : "") + ex;
}

// error.rs:73-110
inline LinalgError singular_matrix_with_suggestions(const std::string &operation, std::pair<std::size_t, std::size_t> shape,
                                                    std::optional<double> condition_number)
{
    std::string base = "Matrix is singular during " + operation + " operation";
    std::string shape_info = "Matrix shape: " + shape_str(shape.first, shape.second);
    std::vector<std::string> s = {
        "Consider the following regularization approaches:",
        "1. Ridge regularization: Add \xCE\xBBI to the matrix (small positive \xCE\xBB)",
        "2. Pseudo-inverse: Use SVD-based pseudo-inverse for rank-deficient matrices",
        "3. Truncated SVD: Remove small singular values below a threshold"};
    if (condition_number) {
        s.push_back("4. Condition number: " + fmt_e2(*condition_number) + " (>1e12 indicates ill-conditioning)");
        if (*condition_number > 1e12) s.push_back("5. Use extended precision arithmetic for better numerical stability");
    }
    s.push_back("6. Check input data for linear dependencies or scaling issues");
    s.push_back("7. Use iterative refinement for improved accuracy");
    std::string joined;
    for (std::size_t i = 0; i < s.size(); ++i) joined += (i ? "\n" : "") + s[i];
    return LinalgError(LinalgError::Kind::SingularMatrixError, base + "\n" + shape_info + "\n" + joined);
}

// error.rs:112-150
inline LinalgError non_positive_definite_with_suggestions(const std::string &operation,
                                                          std::pair<std::size_t, std::size_t> shape,
                                                          std::optional<std::size_t> negative_eigenvalues)
{
    std::string base = "Matrix is not positive definite during " + operation + " operation";
    std::string shape_info = "Matrix shape: " + shape_str(shape.first, shape.second);
    std::vector<std::string> s = {
        "Consider the following regularization approaches:",
        "1. Diagonal regularization: Add \xCE\xBBI where \xCE\xBB > |most negative eigenvalue|",
        "2. Modified Cholesky: Use algorithms that ensure positive definiteness",
        "3. Eigenvalue clipping: Replace negative eigenvalues with small positive values",
        "4. Use LDL decomposition instead of Cholesky for indefinite matrices"};
    if (negative_eigenvalues)
        s.push_back("5. Found " + std::to_string(*negative_eigenvalues) +
                    " negative eigenvalue(s) - consider spectral regularization");
    s.push_back("6. Check if matrix is symmetric (required for Cholesky)");
    s.push_back("7. Use pivoted Cholesky for rank-deficient positive semidefinite matrices");
    s.push_back("8. Consider using QR or LU decomposition for non-symmetric matrices");
    std::string joined;
    for (std::size_t i = 0; i < s.size(); ++i) joined += (i ? "\n" : "") + s[i];
    return LinalgError(LinalgError::Kind::NonPositiveDefiniteError, base + "\n" + shape_info + "\n" + joined);
}

// LinalgResult<T> (error.rs:202)
template <class T>
class Result {
  public:
    Result(T v) : ok_(true), val_(std::move(v)), err_(LinalgError::Kind::ComputationError, "") {}
    Result(LinalgError e) : ok_(false), val_(), err_(std::move(e)) {}
    bool is_ok() const { return ok_; }
    bool is_err() const { return !ok_; }
    const T &value() const { return val_; }
    T &value() { return val_; }
    const LinalgError &error() const { return err_; }

  private:
    bool ok_;
    T val_;
    LinalgError err_;
};

// ------------------------------------------------------------------ arrays
// Strided views with the meaning of ndarray's ArrayView1/ArrayView2.
template <class F>
struct ArrayView1 {
    const F *data = nullptr;
    std::size_t len = 0;
    std::ptrdiff_t stride = 1;
    F operator[](std::size_t i) const { return data[(std::ptrdiff_t)i * stride]; }
    bool is_empty() const { return len == 0; }
};

template <class F>
struct ArrayView2 {
    const F *data = nullptr;
    std::size_t rows = 0, cols = 0;
    std::ptrdiff_t row_stride = 0, col_stride = 1;
    F operator()(std::size_t i, std::size_t j) const
    {
        return data[(std::ptrdiff_t)i * row_stride + (std::ptrdiff_t)j * col_stride];
    }
    std::size_t nrows() const { return rows; }
    std::size_t ncols() const { return cols; }
    bool is_empty() const { return rows == 0 || cols == 0; }
    std::pair<std::size_t, std::size_t> dim() const { return {rows, cols}; }
};

template <class F>
struct Array2 {
    std::size_t rows = 0, cols = 0;
    std::vector<F> data;   // row-major
    Array2() = default;
    Array2(std::size_t r, std::size_t c) : rows(r), cols(c), data(r * c, F(0)) {}
    F &operator()(std::size_t i, std::size_t j) { return data[i * cols + j]; }
    F operator()(std::size_t i, std::size_t j) const { return data[i * cols + j]; }
    ArrayView2<F> view() const { return {data.data(), rows, cols, (std::ptrdiff_t)cols, 1}; }
};

template <class F>
using Array1 = std::vector<F>;

// ndarray's ArrayView3 / Array3 (any strides in the view; the owned array is C order)
template <class F>
struct ArrayView3 {
    const F *data = nullptr;
    std::size_t d0 = 0, d1 = 0, d2 = 0;
    std::ptrdiff_t s0 = 0, s1 = 0, s2 = 1;
    F operator()(std::size_t p, std::size_t i, std::size_t j) const
    {
        return data[(std::ptrdiff_t)p * s0 + (std::ptrdiff_t)i * s1 + (std::ptrdiff_t)j * s2];
    }
    std::tuple<std::size_t, std::size_t, std::size_t> dim() const { return {d0, d1, d2}; }
};

template <class F>
struct Array3 {
    std::size_t d0 = 0, d1 = 0, d2 = 0;
    std::vector<F> data;
    Array3() = default;
    Array3(std::size_t a, std::size_t b, std::size_t c) : d0(a), d1(b), d2(c), data(a * b * c, F(0)) {}
    ArrayView3<F> view() const
    {
        return {data.data(), d0, d1, d2, (std::ptrdiff_t)(d1 * d2), (std::ptrdiff_t)d2, 1};
    }
};

template <class F>
ArrayView1<F> view1(const std::vector<F> &v)
{
    return {v.data(), v.size(), 1};
}

// ------------------------------------------------------------------ validation (validation.rs)
template <class F>
std::optional<LinalgError> validate_not_empty_matrix(const ArrayView2<F> &m, const std::string &op)
{
    if (m.is_empty()) return LinalgError(LinalgError::Kind::ShapeError, op + " failed: Input matrix cannot be empty");
    return std::nullopt;
}
template <class F>
std::optional<LinalgError> validate_not_empty_vector(const ArrayView1<F> &v, const std::string &op)
{
    if (v.is_empty()) return LinalgError(LinalgError::Kind::ShapeError, op + " failed: Input vector cannot be empty");
    return std::nullopt;
}
// validation.rs:69-81 (the Rust format string carries literal "\n" sequences)
template <class F>
std::optional<LinalgError> validate_square_matrix(const ArrayView2<F> &m, const std::string &op)
{
    if (m.nrows() != m.ncols())
        return LinalgError(LinalgError::Kind::ShapeError,
                           op + " failed: Matrix must be square\\nMatrix shape: " + shape_str(m.nrows(), m.ncols()) +
                               "\\nExpected: Square matrix (n\xC3\x97n)");
    return std::nullopt;
}
template <class F>
std::optional<LinalgError> validate_matrix_vector_dimensions(const ArrayView2<F> &m, const ArrayView1<F> &v,
                                                             const std::string &op)
{
    if (m.nrows() != v.len)
        return LinalgError(LinalgError::Kind::ShapeError,
                           op + " failed: Matrix and vector dimensions must match\\nMatrix shape: " +
                               shape_str(m.nrows(), m.ncols()) + "\\nVector shape: " + std::to_string(v.len) +
                               "\\nExpected: Vector length = " + std::to_string(m.nrows()));
    return std::nullopt;
}
template <class F>
std::optional<LinalgError> validate_matrix_matrix_dimensions(const ArrayView2<F> &a, const ArrayView2<F> &b,
                                                             const std::string &op, bool require_same_rows)
{
    if (require_same_rows && a.nrows() != b.nrows())
        return LinalgError(LinalgError::Kind::ShapeError,
                           op + " failed: Matrix rows must match\\nFirst matrix shape: " + shape_str(a.nrows(), a.ncols()) +
                               "\\nSecond matrix shape: " + shape_str(b.nrows(), b.ncols()) +
                               "\\nExpected: Same number of rows");
    return std::nullopt;
}
template <class F>
std::optional<LinalgError> validate_finite_matrix(const ArrayView2<F> &m, const std::string &op)
{
    for (std::size_t i = 0; i < m.nrows(); ++i)
        for (std::size_t j = 0; j < m.ncols(); ++j)
            if (!std::isfinite(m(i, j)))
                return LinalgError(LinalgError::Kind::InvalidInputError, op + " failed: Matrix contains non-finite values");
    return std::nullopt;
}
template <class F>
std::optional<LinalgError> validate_finite_vector(const ArrayView1<F> &v, const std::string &op)
{
    for (std::size_t i = 0; i < v.len; ++i)
        if (!std::isfinite(v[i]))
            return LinalgError(LinalgError::Kind::InvalidInputError, op + " failed: Vector contains non-finite values");
    return std::nullopt;
}
// validation.rs:213-228
template <class F>
std::optional<LinalgError> validate_linear_system(const ArrayView2<F> &m, const ArrayView1<F> &v, const std::string &op)
{
    if (auto e = validate_not_empty_matrix(m, op)) return e;
    if (auto e = validate_not_empty_vector(v, op)) return e;
    if (auto e = validate_square_matrix(m, op)) return e;
    if (auto e = validate_matrix_vector_dimensions(m, v, op)) return e;
    if (auto e = validate_finite_matrix(m, op)) return e;
    if (auto e = validate_finite_vector(v, op)) return e;
    return std::nullopt;
}
// validation.rs:280-294
template <class F>
std::optional<LinalgError> validate_decomposition(const ArrayView2<F> &m, const std::string &op, bool require_square)
{
    if (auto e = validate_not_empty_matrix(m, op)) return e;
    if (require_square)
        if (auto e = validate_square_matrix(m, op)) return e;
    if (auto e = validate_finite_matrix(m, op)) return e;
    return std::nullopt;
}
// validation.rs:314-329
template <class F>
std::optional<LinalgError> validate_multiple_linear_systems(const ArrayView2<F> &a, const ArrayView2<F> &b,
                                                            const std::string &op)
{
    if (auto e = validate_not_empty_matrix(a, op)) return e;
    if (auto e = validate_not_empty_matrix(b, op)) return e;
    if (auto e = validate_square_matrix(a, op)) return e;
    if (auto e = validate_matrix_matrix_dimensions(a, b, op, true)) return e;
    if (auto e = validate_finite_matrix(a, op)) return e;
    if (auto e = validate_finite_matrix(b, op)) return e;
    return std::nullopt;
}

// ------------------------------------------------------------------ GPU dispatch
namespace detail {
template <class F>
constexpr bool supported = std::is_same<F, double>::value || std::is_same<F, float>::value;

// row-major contiguous copy of a strided view (ndarray's as_standard_layout)
template <class F>
std::vector<F> standard(const ArrayView2<F> &a)
{
    std::vector<F> out(a.rows * a.cols);
    for (std::size_t i = 0; i < a.rows; ++i)
        for (std::size_t j = 0; j < a.cols; ++j) out[i * a.cols + j] = a(i, j);
    return out;
}

inline LinalgError gpu_failure(int rc)
{
    return LinalgError(LinalgError::Kind::ComputationError,
                       std::string("GPU solve failed (") + std::to_string(rc) + "): " + scirs_gpu_last_error());
}

inline void set_workers(std::optional<std::size_t> workers)
{
    // solve.rs:79-81: the only effect of `workers` is this environment variable
    if (workers) setenv("OMP_NUM_THREADS", std::to_string(*workers).c_str(), 1);
}

template <class F>
int getrf(int m, int n, F *a, int64_t *perm, scirs_solve_info *info)
{
    if constexpr (std::is_same<F, double>::value) return scirs_gpu_dgetrf(m, n, a, n, perm, info);
    else return scirs_gpu_sgetrf(m, n, a, n, perm, info);
}
template <class F>
int gesv(int n, int nrhs, const F *a, const F *b, F *x, scirs_solve_info *info)
{
    if constexpr (std::is_same<F, double>::value) return scirs_gpu_dgesv(n, nrhs, a, n, b, nrhs, x, nrhs, info);
    else return scirs_gpu_sgesv(n, nrhs, a, n, b, nrhs, x, nrhs, info);
}
template <class F>
int potrf(int n, F *a, scirs_solve_info *info)
{
    if constexpr (std::is_same<F, double>::value) return scirs_gpu_dpotrf(n, a, n, info);
    else return scirs_gpu_spotrf(n, a, n, info);
}
template <class F>
int trsv(int n, const F *a, const F *b, F *x, bool lower, bool unit_diagonal, scirs_solve_info *info)
{
    if constexpr (std::is_same<F, double>::value)
        return scirs_gpu_dtrsv(n, 1, a, n, b, 1, x, 1, lower, unit_diagonal, info);
    else return scirs_gpu_strsv(n, 1, a, n, b, 1, x, 1, lower, unit_diagonal, info);
}
template <class F>
int posv(int n, int nrhs, const F *a, const F *b, F *x, scirs_solve_info *info)
{
    if constexpr (std::is_same<F, double>::value) return scirs_gpu_dposv(n, nrhs, a, n, b, nrhs, x, nrhs, info);
    else return scirs_gpu_sposv(n, nrhs, a, n, b, nrhs, x, nrhs, info);
}
} // namespace detail

template <class F>
Result<Array2<F>> solve_multiple(const ArrayView2<F> &a, const ArrayView2<F> &b, std::optional<std::size_t> workers);
template <class F>
Result<Array2<F>> inv(const ArrayView2<F> &a, std::optional<std::size_t> workers);

// decomposition.rs:108-183: (P, L, U) with P A = L U
template <class F>
Result<std::tuple<Array2<F>, Array2<F>, Array2<F>>> lu(const ArrayView2<F> &a, std::optional<std::size_t> workers)
{
    static_assert(detail::supported<F>, "the GPU path handles f32 and f64");
    if (a.is_empty())
        return LinalgError(LinalgError::Kind::ShapeError, "LU decomposition failed: Input matrix cannot be empty");
    for (std::size_t i = 0; i < a.rows; ++i)
        for (std::size_t j = 0; j < a.cols; ++j)
            if (!std::isfinite(a(i, j)))
                return LinalgError(LinalgError::Kind::InvalidInputError,
                                   "LU decomposition failed: Matrix contains non-finite values");
    detail::set_workers(workers);
    const std::size_t n = a.rows, m = a.cols, k = n < m ? n : m;
    std::vector<F> packed = detail::standard(a);
    std::vector<int64_t> piv(n);
    scirs_solve_info info;
    int rc = detail::getrf<F>((int)n, (int)m, packed.data(), piv.data(), &info);
    if (rc > 0) return singular_matrix_with_suggestions("LU decomposition", a.dim(), std::nullopt);
    if (rc < 0) return detail::gpu_failure(rc);
    Array2<F> p(n, n), l(n, k), u(k, m);
    for (std::size_t i = 0; i < n; ++i) p(i, (std::size_t)piv[i]) = F(1);
    for (std::size_t i = 0; i < n; ++i) {
        for (std::size_t j = 0; j < (i < m ? i : m); ++j) l(i, j) = packed[i * m + j];
        if (i < m) l(i, i) = F(1);
    }
    for (std::size_t i = 0; i < k; ++i)
        for (std::size_t j = i; j < m; ++j) u(i, j) = packed[i * m + j];
    return std::make_tuple(std::move(p), std::move(l), std::move(u));
}

// decomposition.rs:44-81
template <class F>
Result<Array2<F>> cholesky(const ArrayView2<F> &a, std::optional<std::size_t> workers)
{
    static_assert(detail::supported<F>, "the GPU path handles f32 and f64");
    if (auto e = validate_decomposition(a, "Cholesky decomposition", true)) return *e;
    detail::set_workers(workers);
    const std::size_t n = a.rows;
    Array2<F> l(n, n);
    l.data = detail::standard(a);
    scirs_solve_info info;
    int rc = detail::potrf<F>((int)n, l.data.data(), &info);
    if (rc > 0) return non_positive_definite_with_suggestions("Cholesky decomposition", a.dim(), std::nullopt);
    if (rc < 0) return detail::gpu_failure(rc);
    return l;
}

// solve.rs:139-202
template <class F>
Result<Array1<F>> solve_triangular(const ArrayView2<F> &a, const ArrayView1<F> &b, bool lower, bool unit_diagonal)
{
    const std::string op = "Triangular system solve";
    if (auto e = validate_not_empty_matrix(a, op)) return *e;
    if (auto e = validate_not_empty_vector(b, op)) return *e;
    if (auto e = validate_square_matrix(a, op)) return *e;
    if (auto e = validate_matrix_vector_dimensions(a, b, op)) return *e;
    if (auto e = validate_finite_matrix(a, op)) return *e;
    if (auto e = validate_finite_vector(b, op)) return *e;
    static_assert(detail::supported<F>, "the GPU path handles f32 and f64");
    const std::size_t n = a.rows;
    // the substitution runs in F, as the reference's loop does
    std::vector<F> ad = detail::standard(a), bd(n), xd(n);
    for (std::size_t i = 0; i < n; ++i) bd[i] = b[i];
    // The reference stops at the first diagonal with |a_ii| < F::epsilon() in
    // substitution order (solve.rs:166-173, 186-193); only the error comes back,
    // so the test runs here, before any device work.
    if (!unit_diagonal)
        for (std::size_t k = 0; k < n; ++k) {
            std::size_t i = lower ? k : n - 1 - k;
            if (std::fabs(a(i, i)) < std::numeric_limits<F>::epsilon())
                return singular_matrix_with_suggestions(lower ? "triangular system solve (forward substitution)"
                                                              : "triangular system solve (back substitution)",
                                                        a.dim(), 1e16);
        }
    scirs_solve_info info;
    int rc = detail::trsv<F>((int)n, ad.data(), bd.data(), xd.data(), lower, unit_diagonal, &info);
    if (rc > 0)
        return singular_matrix_with_suggestions(lower ? "triangular system solve (forward substitution)"
                                                      : "triangular system solve (back substitution)",
                                                a.dim(), 1e16);
    if (rc < 0) return detail::gpu_failure(rc);
    Array1<F> x(n);
    for (std::size_t i = 0; i < n; ++i) x[i] = xd[i];
    return x;
}

// solve.rs:54-111
template <class F>
Result<Array1<F>> solve(const ArrayView2<F> &a, const ArrayView1<F> &b, std::optional<std::size_t> workers)
{
    static_assert(detail::supported<F>, "the GPU path handles f32 and f64");
    if (auto e = validate_linear_system(a, b, "Linear system solve")) return *e;
    const std::size_t n = a.rows;
    if (n <= 4) {
        // solve.rs:65-76: small systems go through the explicit inverse
        auto ai = inv(a, std::nullopt);
        if (ai.is_err()) return ai.error();
        Array1<F> x(n, F(0));
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = 0; j < n; ++j) x[i] += ai.value()(i, j) * b[j];
        return x;
    }
    detail::set_workers(workers);
    std::vector<F> ad = detail::standard(a), bd(n), xd(n);
    for (std::size_t i = 0; i < n; ++i) bd[i] = b[i];
    scirs_solve_info info;
    int rc = detail::gesv<F>((int)n, 1, ad.data(), bd.data(), xd.data(), &info);
    if (rc > 0) return singular_matrix_with_suggestions("linear system solve", a.dim(), std::nullopt);
    if (rc < 0) return detail::gpu_failure(rc);
    return xd;
}

// solve.rs:366-424
template <class F>
Result<Array2<F>> solve_multiple(const ArrayView2<F> &a, const ArrayView2<F> &b, std::optional<std::size_t> workers)
{
    static_assert(detail::supported<F>, "the GPU path handles f32 and f64");
    if (auto e = validate_multiple_linear_systems(a, b, "Multiple linear systems solve")) return *e;
    detail::set_workers(workers);
    const std::size_t n = a.rows, nrhs = b.cols;
    std::vector<F> ad = detail::standard(a), bd = detail::standard(b);
    Array2<F> x(a.cols, nrhs);
    scirs_solve_info info;
    int rc = detail::gesv<F>((int)n, (int)nrhs, ad.data(), bd.data(), x.data.data(), &info);
    if (rc > 0) return singular_matrix_with_suggestions("multiple linear systems solve", a.dim(), std::nullopt);
    if (rc < 0) return detail::gpu_failure(rc);
    return x;
}

// basic.rs:29-93 (the n > 3 branch through lu)
template <class F>
Result<F> det(const ArrayView2<F> &a, std::optional<std::size_t> workers)
{
    if (a.nrows() != a.ncols())
        return LinalgError(LinalgError::Kind::ShapeError,
                           "Determinant computation failed: Matrix must be square\nMatrix shape: " +
                               shape_str(a.nrows(), a.ncols()) + "\nExpected: Square matrix (n\xC3\x97n)");
    switch (a.nrows()) {
    case 0: return F(1);
    case 1: return a(0, 0);
    case 2: return a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0);
    case 3:
        return a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) - a(0, 1) * (a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0)) +
               a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0));
    default: break;
    }
    auto r = lu(a, workers);
    if (r.is_err()) {
        if (r.error().kind() == LinalgError::Kind::SingularMatrixError) return F(0);
        return r.error();
    }
    const auto &p = std::get<0>(r.value());
    const auto &u = std::get<2>(r.value());
    F det_u = F(1);
    for (std::size_t i = 0; i < u.rows; ++i) det_u *= u(i, i);
    std::size_t swaps = 0;
    for (std::size_t i = 0; i < p.rows; ++i)
        for (std::size_t j = 0; j < i; ++j)
            if (p(i, j) == F(1)) ++swaps;
    return swaps % 2 == 0 ? det_u : -det_u;
}

// basic.rs:117-186
template <class F>
Result<Array2<F>> inv(const ArrayView2<F> &a, std::optional<std::size_t> workers)
{
    if (a.nrows() != a.ncols())
        return LinalgError(LinalgError::Kind::ShapeError,
                           "Matrix inverse computation failed: Matrix must be square\nMatrix shape: " +
                               shape_str(a.nrows(), a.ncols()) + "\nExpected: Square matrix (n\xC3\x97n)");
    if (a.nrows() == 2) {
        auto d = det(a, workers);
        if (d.is_err()) return d.error();
        const F det_val = d.value();
        if (std::fabs(det_val) < std::numeric_limits<F>::epsilon()) {
            const F norm_a = std::sqrt(a(0, 0) * a(0, 0) + a(0, 1) * a(0, 1) + a(1, 0) * a(1, 0) + a(1, 1) * a(1, 1));
            std::optional<double> cond;
            if (std::fabs(det_val) > F(0)) cond = (double)(norm_a / std::fabs(det_val));
            return singular_matrix_with_suggestions("matrix inverse", a.dim(), cond);
        }
        const F inv_det = F(1) / det_val;
        Array2<F> r(2, 2);
        r(0, 0) = a(1, 1) * inv_det;
        r(0, 1) = -a(0, 1) * inv_det;
        r(1, 0) = -a(1, 0) * inv_det;
        r(1, 1) = a(0, 0) * inv_det;
        return r;
    }
    const std::size_t n = a.nrows();
    Array2<F> eye(n, n);
    for (std::size_t i = 0; i < n; ++i) eye(i, i) = F(1);
    auto r = solve_multiple(a, eye.view(), workers);
    if (r.is_err() && r.error().kind() == LinalgError::Kind::SingularMatrixError)
        return singular_matrix_with_suggestions("matrix inverse via solve", a.dim(), std::nullopt);
    return r;
}

// ------------------------------------------------------------------ mixed precision
// `{:?}` of an ndarray shape slice: "[r, c]"
inline std::string shape_dbg(std::size_t r, std::size_t c)
{
    return "[" + std::to_string(r) + ", " + std::to_string(c) + "]";
}

// mixed_precision/adaptive.rs:49-151: solve A x = b with the elimination done
// in precision H and the result returned in C.  The reference eliminates on
// [A|b] with the first-max partial pivot and the absolute test
// max |a_ji| < H::epsilon(); the GPU LU in precision H follows the same rule, so
// a singular system fails the same way (SingularMatrixError "Matrix is
// singular").
template <class A, class B, class C, class H>
Result<Array1<C>> mixed_precision_solve(const ArrayView2<A> &a, const ArrayView1<B> &b)
{
    static_assert(detail::supported<H>, "the GPU path computes in f32 or f64");
    if (a.rows != a.cols)
        return LinalgError(LinalgError::Kind::ShapeError,
                           "Matrix must be square, got shape " + shape_dbg(a.rows, a.cols));
    if (a.rows != b.len)
        return LinalgError(LinalgError::Kind::ShapeError, "Matrix rows (" + std::to_string(a.rows) +
                                                              ") must match vector length (" +
                                                              std::to_string(b.len) + ")");
    const std::size_t n = a.rows;
    Array1<C> out(n, C(0));
    if (n == 0) return out;
    std::vector<H> ah(n * n), bh(n), xh(n);
    for (std::size_t i = 0; i < n; ++i) {
        for (std::size_t j = 0; j < n; ++j) ah[i * n + j] = static_cast<H>(a(i, j));
        bh[i] = static_cast<H>(b[i]);
    }
    scirs_solve_info info;
    int rc = detail::gesv<H>((int)n, 1, ah.data(), bh.data(), xh.data(), &info);
    if (rc > 0) return LinalgError(LinalgError::Kind::SingularMatrixError, "Matrix is singular");
    if (rc < 0) return detail::gpu_failure(rc);
    for (std::size_t i = 0; i < n; ++i) out[i] = static_cast<C>(xh[i]);
    return out;
}

// mixed_precision/adaptive.rs:282-378: solve in working precision W, refine in
// higher precision H, return C.  The reference re-factorises in W for every
// correction and stops on the absolute ||r||_inf < tol (default 1e-8) or after
// max_iter (default 10) corrections.  Here, with H = f64 and W = f32, the
// system is factored once in FP32 on the GPU and refined with FP64 residuals
// until the normwise backward error reaches n * 2^-53 (scirs_gpu_dgesv_ir): the
// FP64-accurate solution, at least as accurate as the reference's stopping rule
// gives; `tol` is not used on that path.  Errors of the working-precision
// solve are those of `solve` (solve.rs:54-111).
template <class A, class B, class C, class H, class W>
Result<Array1<C>> iterative_refinement_solve(const ArrayView2<A> &a, const ArrayView1<B> &b,
                                             std::optional<std::size_t> max_iter, std::optional<H> tol)
{
    static_assert(detail::supported<H> && detail::supported<W>, "the GPU path computes in f32 or f64");
    const std::size_t iters = max_iter.value_or(10);
    if (a.rows != a.cols)
        return LinalgError(LinalgError::Kind::ShapeError,
                           "Expected square matrix for iterative refinement, got " + shape_dbg(a.rows, a.cols));
    if (a.rows != b.len)
        return LinalgError(LinalgError::Kind::ShapeError, "Matrix shape " + shape_dbg(a.rows, a.cols) +
                                                              " is incompatible with vector length " +
                                                              std::to_string(b.len));
    const std::size_t n = a.rows;
    Array1<C> out(n, C(0));
    if constexpr (std::is_same<H, double>::value && std::is_same<W, float>::value) {
        if (n > 4) {
            std::vector<double> ah(n * n), bh(n), xh(n);
            for (std::size_t i = 0; i < n; ++i) {
                for (std::size_t j = 0; j < n; ++j) ah[i * n + j] = static_cast<double>(a(i, j));
                bh[i] = static_cast<double>(b[i]);
            }
            scirs_solve_info info;
            int rc = scirs_gpu_dgesv_ir((int)n, 1, ah.data(), (int64_t)n, bh.data(), 1, xh.data(), 1,
                                        SCIRS_MODE_AUTO, (int)iters, 0, 0.0, &info);
            if (rc > 0) return singular_matrix_with_suggestions("linear system solve", a.dim(), std::nullopt);
            if (rc < 0) return detail::gpu_failure(rc);
            for (std::size_t i = 0; i < n; ++i) out[i] = static_cast<C>(xh[i]);
            return out;
        }
    }
    // Otherwise the reference's loop itself (n <= 4, or other precisions):
    // solves in W through `solve` (the GPU for n > 4, the inverse branch below),
    // residuals in H on the host (O(n^2)), stop on ||r||_inf < tol.
    const H tl = tol.value_or(static_cast<H>(1e-8));
    std::vector<W> aw(n * n), bw(n);
    for (std::size_t i = 0; i < n; ++i) {
        for (std::size_t j = 0; j < n; ++j) aw[i * n + j] = static_cast<W>(a(i, j));
        bw[i] = static_cast<W>(b[i]);
    }
    ArrayView2<W> av{aw.data(), n, n, (std::ptrdiff_t)n, 1};
    auto x = solve<W>(av, view1(bw), std::nullopt);
    if (x.is_err()) return x.error();
    std::vector<H> xh(n), r(n);
    for (std::size_t i = 0; i < n; ++i) xh[i] = static_cast<H>(x.value()[i]);
    for (std::size_t it = 0; it < iters; ++it) {
        H rn = H(0);
        for (std::size_t i = 0; i < n; ++i) {
            H ax = H(0);
            for (std::size_t j = 0; j < n; ++j) ax += static_cast<H>(a(i, j)) * xh[j];
            r[i] = static_cast<H>(b[i]) - ax;
            rn = std::max(rn, std::abs(r[i]));
        }
        if (rn < tl) break;
        std::vector<W> rw(n);
        for (std::size_t i = 0; i < n; ++i) rw[i] = static_cast<W>(r[i]);
        auto dx = solve<W>(av, view1(rw), std::nullopt);
        if (dx.is_err()) return dx.error();
        for (std::size_t i = 0; i < n; ++i) xh[i] += static_cast<H>(dx.value()[i]);
    }
    for (std::size_t i = 0; i < n; ++i) out[i] = static_cast<C>(xh[i]);
    return out;
}

// ------------------------------------------------------------------ compat
// decomposition.rs:208-249 (qr over lapack.rs:171-253), f64 on the GPU.
template <class F>
Result<std::pair<Array2<F>, Array2<F>>> qr(const ArrayView2<F> &a, std::optional<std::size_t> workers)
{
    static_assert(std::is_same<F, double>::value, "the GPU QR handles f64");
    if (a.is_empty())
        return LinalgError(LinalgError::Kind::ShapeError, "QR decomposition failed: Input matrix cannot be empty");
    for (std::size_t i = 0; i < a.rows; ++i)
        for (std::size_t j = 0; j < a.cols; ++j)
            if (!std::isfinite(a(i, j)))
                return LinalgError(LinalgError::Kind::InvalidInputError,
                                   "QR decomposition failed: Matrix contains non-finite values");
    detail::set_workers(workers);
    // lapack.rs:185-190
    if (a.rows < a.cols)
        return LinalgError(LinalgError::Kind::ComputationError, "QR decomposition requires rows >= columns");
    std::vector<F> ad = detail::standard(a);
    Array2<F> q(a.rows, a.rows), r(a.rows, a.cols);
    scirs_solve_info info;
    int rc = scirs_gpu_dqr((int)a.rows, (int)a.cols, ad.data(), (int64_t)a.cols, q.data.data(), (int64_t)a.rows,
                           r.data.data(), (int64_t)a.cols, &info);
    if (rc < 0) return detail::gpu_failure(rc);
    return std::make_pair(std::move(q), std::move(r));
}

// solve.rs:17-26 (LstsqResult)
template <class F>
struct LstsqResult {
    Array1<F> x;
    F residuals;
    std::size_t rank;
    Array1<F> s;
};

// solve.rs:234-290: the rows >= cols (QR) branch runs on the GPU; the rows <
// cols branch is the reference's SVD path, which this library does not
// provide (the shim keeps the reference's own code for it).
template <class F>
Result<LstsqResult<F>> lstsq(const ArrayView2<F> &a, const ArrayView1<F> &b, std::optional<std::size_t> workers)
{
    static_assert(std::is_same<F, double>::value, "the GPU least-squares solve handles f64");
    if (auto e = validate_not_empty_matrix(a, "Least squares solve")) return *e;
    if (auto e = validate_not_empty_vector(b, "Least squares solve")) return *e;
    if (auto e = validate_matrix_vector_dimensions(a, b, "Least squares solve")) return *e;
    if (auto e = validate_finite_matrix(a, "Least squares solve")) return *e;
    if (auto e = validate_finite_vector(b, "Least squares solve")) return *e;
    detail::set_workers(workers);
    if (a.rows < a.cols)
        return LinalgError(LinalgError::Kind::NotImplementedError,
                           "lstsq with rows < columns (the SVD branch) is not on the GPU path");
    std::vector<F> ad = detail::standard(a), bd(a.rows);
    for (std::size_t i = 0; i < a.rows; ++i) bd[i] = b[i];
    LstsqResult<F> out{Array1<F>(a.cols), F(0), a.cols, Array1<F>()};
    scirs_solve_info info;
    int rc = scirs_gpu_dlstsq((int)a.rows, (int)a.cols, ad.data(), (int64_t)a.cols, bd.data(), out.x.data(),
                              &out.residuals, &info);
    if (rc > 0)
        return singular_matrix_with_suggestions("triangular system solve (back substitution)", {a.cols, a.cols}, 1e16);
    if (rc < 0) return detail::gpu_failure(rc);
    return out;
}

// ------------------------------------------------------------------ Krylov
// solvers/iterative.rs:13-50 (IterativeSolverOptions, IterativeSolverResult)
template <class A>
struct IterativeSolverOptions {
    std::size_t max_iterations = 1000;
    A tolerance = A(1e-10);
    bool verbose = false;
    std::optional<std::size_t> restart;
};

template <class A>
struct IterativeSolverResult {
    Array1<A> solution;
    std::size_t iterations;
    A residual_norm;
    bool converged;
};

// solvers/iterative.rs:317-450 (gmres), f64 on the GPU
template <class A>
Result<IterativeSolverResult<A>> gmres(const ArrayView2<A> &a, const ArrayView1<A> &b, const ArrayView1<A> *x0,
                                       const IterativeSolverOptions<A> &options)
{
    static_assert(std::is_same<A, double>::value, "the GPU GMRES handles f64");
    const std::size_t n = a.rows;
    if (a.cols != n) return LinalgError(LinalgError::Kind::ShapeError, "Matrix A must be square");
    if (b.len != n)
        return LinalgError(LinalgError::Kind::ShapeError, "Vector b must have same length as matrix dimension");
    std::vector<A> x0d;
    if (x0) {
        if (x0->len != n)
            return LinalgError(LinalgError::Kind::ShapeError, "Initial guess x0 must have same length as b");
        x0d.resize(n);
        for (std::size_t i = 0; i < n; ++i) x0d[i] = (*x0)[i];
    }
    if (options.restart && *options.restart == 0)
        // the reference computes restart.min(n) = 0 and then divides by it
        // (iterative.rs:361-366), a panic; here it is an error
        return LinalgError(LinalgError::Kind::InvalidInputError, "GMRES restart must be positive");
    IterativeSolverResult<A> out{Array1<A>(n), 0, A(0), false};
    if (n == 0) {
        // the reference divides by restart = min(n, 30) = 0 here and panics;
        // this returns the empty solution instead
        out.converged = A(0) < options.tolerance;
        return out;
    }
    std::vector<A> ad = detail::standard(a), bd(n);
    for (std::size_t i = 0; i < n; ++i) bd[i] = b[i];
    int it = 0, conv = 0;
    double res = 0.0;
    int rc = scirs_gpu_dgmres((int)n, ad.data(), (int64_t)n, bd.data(), x0 ? x0d.data() : nullptr,
                              (int)options.max_iterations, (double)options.tolerance,
                              options.restart ? (int)*options.restart : 0, out.solution.data(), &it, &res, &conv,
                              nullptr);
    if (rc > 0)
        return LinalgError(LinalgError::Kind::SingularMatrixError, "Matrix is singular in GMRES least squares solve");
    if (rc < 0) return detail::gpu_failure(rc);
    // options.verbose is accepted and not acted on: the reference prints one
    // "GMRES outer iteration {}: residual = {}" line per restart cycle
    // (iterative.rs:424-429), which the GPU loop does not report back
    out.iterations = (std::size_t)it;
    out.residual_norm = (A)res;
    out.converged = conv != 0;
    return out;
}

// preconditioners.rs:179-196 (PreconditionerOp).  The GPU solve below knows
// two preconditioners by type; any other PreconditionerOp is refused (its
// apply is host code the device iteration cannot call per step).
template <class F>
struct PreconditionerOp {
    virtual ~PreconditionerOp() = default;
    virtual Result<Array1<F>> apply(const ArrayView1<F> &x) const = 0;
    virtual Result<Array1<F>> apply_transpose(const ArrayView1<F> &x) const { return apply(x); }
    virtual std::size_t size() const = 0;
    virtual bool is_symmetric() const { return true; }
    // which device preconditioner stands in for this one (SCIRS_PRECOND_*), -1: none
    virtual int gpu_kind() const { return -1; }
};

// preconditioners.rs:198-245: 1/a_ii, or 1 where |a_ii| < eps
template <class F>
class DiagonalPreconditioner : public PreconditionerOp<F> {
  public:
    static Result<DiagonalPreconditioner<F>> create(const ArrayView2<F> &matrix)
    {
        if (matrix.rows != matrix.cols)
            return LinalgError(LinalgError::Kind::ShapeError, "Diagonal preconditioner requires square matrix");
        DiagonalPreconditioner<F> p;
        p.inverse_diagonal_ = Array1<F>(matrix.rows);
        for (std::size_t i = 0; i < matrix.rows; ++i) {
            const F d = matrix(i, i);
            p.inverse_diagonal_[i] = std::fabs(d) < std::numeric_limits<F>::epsilon() ? F(1) : F(1) / d;
        }
        return p;
    }
    Result<Array1<F>> apply(const ArrayView1<F> &x) const override
    {
        Array1<F> y(x.len);
        for (std::size_t i = 0; i < x.len; ++i) y[i] = inverse_diagonal_[i] * x[i];
        return y;
    }
    std::size_t size() const override { return inverse_diagonal_.size(); }
    int gpu_kind() const override { return SCIRS_PRECOND_DIAGONAL; }

  private:
    Array1<F> inverse_diagonal_;
};

// No counterpart in the reference: the mixed-precision LU of the matrix (FP32
// tournament-pivoted panel, FP32-accurate update) as a preconditioner, built
// on the GPU inside preconditioned_gmres; apply() on the host is not offered.
template <class F>
class LUPreconditioner : public PreconditionerOp<F> {
  public:
    explicit LUPreconditioner(std::size_t n) : n_(n) {}
    Result<Array1<F>> apply(const ArrayView1<F> &) const override
    {
        return LinalgError(LinalgError::Kind::NotImplementedError,
                           "LUPreconditioner is applied on the GPU inside preconditioned_gmres");
    }
    std::size_t size() const override { return n_; }
    bool is_symmetric() const override { return false; }
    int gpu_kind() const override { return SCIRS_PRECOND_LU; }

  private:
    std::size_t n_;
};

// preconditioners.rs:934-1010 (preconditioned_gmres), f64 on the GPU.  The
// reference's body never solves its least-squares problem (x += 0.1 tol v_0);
// this is the real right-preconditioned restarted GMRES with the gmres
// iteration of solvers/iterative.rs:317-450, so the result solves A x = rhs.
// Shape check and text as preconditioners.rs:947-952; restart 0 (a division
// by zero in the reference) is an InvalidInputError.
template <class F>
Result<Array1<F>> preconditioned_gmres(const ArrayView2<F> &matrix, const ArrayView1<F> &rhs,
                                       const PreconditionerOp<F> &preconditioner, std::size_t max_iterations,
                                       F tolerance, std::size_t restart, const ArrayView1<F> *initial_guess)
{
    static_assert(std::is_same<F, double>::value, "the GPU GMRES handles f64");
    const std::size_t n = matrix.rows;
    if (matrix.cols != n || rhs.len != n)
        return LinalgError(LinalgError::Kind::ShapeError, "Matrix and RHS dimensions must be compatible");
    if (restart == 0) return LinalgError(LinalgError::Kind::InvalidInputError, "GMRES restart must be positive");
    const int kind = preconditioner.gpu_kind();
    if (kind < 0)
        return LinalgError(LinalgError::Kind::NotImplementedError,
                           "preconditioned_gmres on the GPU takes DiagonalPreconditioner or LUPreconditioner");
    Array1<F> x(n);
    if (n == 0) return x;
    std::vector<F> x0d;
    if (initial_guess) {
        if (initial_guess->len != n)
            return LinalgError(LinalgError::Kind::ShapeError, "Matrix and RHS dimensions must be compatible");
        x0d.resize(n);
        for (std::size_t i = 0; i < n; ++i) x0d[i] = (*initial_guess)[i];
    }
    std::vector<F> ad = detail::standard(matrix), bd(n);
    for (std::size_t i = 0; i < n; ++i) bd[i] = rhs[i];
    int it = 0, conv = 0;
    double res = 0.0;
    int rc = scirs_gpu_dgmres_precond((int)n, ad.data(), (int64_t)n, bd.data(), initial_guess ? x0d.data() : nullptr,
                                      (int)max_iterations, (double)tolerance, (int)restart, kind, nullptr, nullptr,
                                      x.data(), &it, &res, &conv, nullptr);
    if (rc > 0)
        return LinalgError(LinalgError::Kind::SingularMatrixError, "Matrix is singular in GMRES least squares solve");
    if (rc < 0) return detail::gpu_failure(rc);
    return x;
}

// ------------------------------------------------------------------ batch (batch/mod.rs)
namespace detail {
template <class F>
std::vector<F> standard3(const ArrayView3<F> &a)
{
    std::vector<F> out(a.d0 * a.d1 * a.d2);
    std::size_t o = 0;
    for (std::size_t p = 0; p < a.d0; ++p)
        for (std::size_t i = 0; i < a.d1; ++i)
            for (std::size_t j = 0; j < a.d2; ++j) out[o++] = a(p, i, j);
    return out;
}
inline int batch_matmul(int nb, int m, int k, int n, const double *a, const double *b, double *c)
{
    return scirs_gpu_dbatch_matmul(nb, m, k, n, a, b, c, nullptr);
}
inline int batch_matmul(int nb, int m, int k, int n, const float *a, const float *b, float *c)
{
    return scirs_gpu_sbatch_matmul(nb, m, k, n, a, b, c, nullptr);
}
inline int batch_matvec(int nb, int m, int n, const double *a, const double *x, double *y)
{
    return scirs_gpu_dbatch_matvec(nb, m, n, a, x, y, nullptr);
}
inline int batch_matvec(int nb, int m, int n, const float *a, const float *x, float *y)
{
    return scirs_gpu_sbatch_matvec(nb, m, n, a, x, y, nullptr);
}
inline int batch_add(int nb, int m, int n, const double *a, const double *v, int axis, double *o)
{
    return scirs_gpu_dbatch_add(nb, m, n, a, v, axis, o, nullptr);
}
inline int batch_add(int nb, int m, int n, const float *a, const float *v, int axis, float *o)
{
    return scirs_gpu_sbatch_add(nb, m, n, a, v, axis, o, nullptr);
}
inline int batch_sum(int nb, int m, int n, const double *a, double *o)
{
    return scirs_gpu_dbatch_sum(nb, m, n, a, o, nullptr);
}
inline int batch_sum(int nb, int m, int n, const float *a, float *o)
{
    return scirs_gpu_sbatch_sum(nb, m, n, a, o, nullptr);
}
} // namespace detail

// batch/mod.rs:62-107
template <class F>
Result<Array3<F>> batch_matmul(const ArrayView3<F> &batch_a, const ArrayView2<F> &b)
{
    static_assert(detail::supported<F>, "the GPU path handles f32 and f64");
    const auto [nb, m, k1] = batch_a.dim();
    if (k1 != b.rows)
        return LinalgError(LinalgError::Kind::ShapeError, "Inner dimensions mismatch for batch_matmul: " +
                                                              std::to_string(k1) + " vs " + std::to_string(b.rows));
    Array3<F> out(nb, m, b.cols);
    std::vector<F> ad = detail::standard3(batch_a), bd = detail::standard(b);
    int rc = detail::batch_matmul((int)nb, (int)m, (int)k1, (int)b.cols, ad.data(), bd.data(), out.data.data());
    if (rc < 0) return detail::gpu_failure(rc);
    return out;
}

// batch/mod.rs:139-172
template <class F>
Result<Array2<F>> batch_matvec(const ArrayView3<F> &batch_a, const ArrayView1<F> &x)
{
    static_assert(detail::supported<F>, "the GPU path handles f32 and f64");
    const auto [nb, m, n] = batch_a.dim();
    if (n != x.len)
        return LinalgError(LinalgError::Kind::ShapeError,
                           "Dimension mismatch for batch_matvec: matrix width " + std::to_string(n) +
                               " does not match vector length " + std::to_string(x.len));
    Array2<F> out(nb, m);
    std::vector<F> ad = detail::standard3(batch_a), xd(x.len);
    for (std::size_t j = 0; j < x.len; ++j) xd[j] = x[j];
    int rc = detail::batch_matvec((int)nb, (int)m, (int)n, ad.data(), xd.data(), out.data.data());
    if (rc < 0) return detail::gpu_failure(rc);
    return out;
}

// batch/mod.rs:220-291
template <class F>
Result<Array3<F>> batch_add(const ArrayView3<F> &batch_a, const ArrayView1<F> &v, std::size_t axis)
{
    static_assert(detail::supported<F>, "the GPU path handles f32 and f64");
    const auto [nb, m, n] = batch_a.dim();
    if (axis == 0) {
        if (v.len != m)
            return LinalgError(LinalgError::Kind::ShapeError,
                               "Column-wise batch_add requires vector length " + std::to_string(m) +
                                   " to match number of rows " + std::to_string(m) + ", got " + std::to_string(v.len));
    } else if (axis == 1) {
        if (v.len != n)
            return LinalgError(LinalgError::Kind::ShapeError,
                               "Row-wise batch_add requires vector length " + std::to_string(n) +
                                   " to match number of columns " + std::to_string(n) + ", got " +
                                   std::to_string(v.len));
    } else {
        return LinalgError(LinalgError::Kind::InvalidInputError,
                           "Invalid axis " + std::to_string(axis) + ": must be 0 (column-wise) or 1 (row-wise)");
    }
    Array3<F> out(nb, m, n);
    std::vector<F> ad = detail::standard3(batch_a), vd(v.len);
    for (std::size_t j = 0; j < v.len; ++j) vd[j] = v[j];
    int rc = detail::batch_add((int)nb, (int)m, (int)n, ad.data(), vd.data(), (int)axis, out.data.data());
    if (rc < 0) return detail::gpu_failure(rc);
    return out;
}

// batch/mod.rs:326-336.  The reference's batch_sum cannot fail and returns the
// array itself; a GPU call can fail (no device, out of memory), so this one
// returns a Result whose value is that array, and a device failure is a
// ComputationError like every other entry point's.
template <class F>
Result<Array2<F>> batch_sum(const ArrayView3<F> &batch_a)
{
    static_assert(detail::supported<F>, "the GPU path handles f32 and f64");
    const auto [nb, m, n] = batch_a.dim();
    Array2<F> out(m, n);
    std::vector<F> ad = detail::standard3(batch_a);
    int rc = detail::batch_sum((int)nb, (int)m, (int)n, ad.data(), out.data.data());
    if (rc < 0) return detail::gpu_failure(rc);
    return out;
}

// compat.rs:461-499 (scipy.linalg.solve).  The reference rejects every
// specialised option with NotImplementedError; here assume_a = "pos" (without
// `lower` / `transposed`) is served by the GPU Cholesky solve, the other
// options keep the reference's error.  overwrite_a / overwrite_b are ignored, as
// in the reference.
template <class F>
Result<Array2<F>> compat_solve(const ArrayView2<F> &a, const ArrayView2<F> &b, bool lower, bool overwrite_a,
                               bool overwrite_b, bool check_finite, std::optional<std::string> assume_a,
                               bool transposed)
{
    static_assert(detail::supported<F>, "the GPU path handles f32 and f64");
    (void)overwrite_a;
    (void)overwrite_b;
    if (check_finite) {
        for (std::size_t i = 0; i < a.rows; ++i)
            for (std::size_t j = 0; j < a.cols; ++j)
                if (!std::isfinite(a(i, j)))
                    return LinalgError(LinalgError::Kind::ValueError, "Matrix contains non-finite values");
        for (std::size_t i = 0; i < b.rows; ++i)
            for (std::size_t j = 0; j < b.cols; ++j)
                if (!std::isfinite(b(i, j)))
                    return LinalgError(LinalgError::Kind::ValueError, "Matrix contains non-finite values");
    }
    const bool pos = assume_a && *assume_a == "pos" && !lower && !transposed;
    if (!pos && (lower || assume_a || transposed))
        return LinalgError(LinalgError::Kind::NotImplementedError, "Specialized solve options not yet supported");
    if (!pos) return solve_multiple(a, b, std::nullopt);
    if (auto e = validate_multiple_linear_systems(a, b, "Multiple linear systems solve")) return *e;
    const std::size_t n = a.rows, nrhs = b.cols;
    // factor and substitutions in F (f32 stays f32)
    std::vector<F> ad = detail::standard(a), bd = detail::standard(b), xd(n * nrhs);
    scirs_solve_info info;
    int rc = detail::posv<F>((int)n, (int)nrhs, ad.data(), bd.data(), xd.data(), &info);
    if (rc > 0) return non_positive_definite_with_suggestions("Cholesky decomposition", a.dim(), std::nullopt);
    if (rc < 0) return detail::gpu_failure(rc);
    Array2<F> x(n, nrhs);
    x.data = std::move(xd);
    return x;
}

} // namespace scirs2_linalg

#endif // SCIRS2_LINALG_HPP
