// Test driver for include/scirs2_linalg.hpp: reads cases from stdin, calls the
// C++ mirror of the scirs2-linalg API, and prints one result line per case.
//
// Input, whitespace separated, one case after another:
//   <op> <type> <flag1> <flag2> <ar> <ac> <ar*ac values> <br> <bc> <br*bc values>
//   op   : solve | solve_multiple | lu | cholesky | solve_triangular | det | inv |
//          mixed_precision_solve | iterative_refinement_solve | compat_solve | qr | lstsq |
//          batch_matmul | batch_matvec | batch_add | batch_sum (flag1 = batch count,
//          A given as (batch*m) x k; flag2 = axis of batch_add)
//   type : f64 | f32
//   flags: lower / unit_diagonal for solve_triangular, ignored otherwise
//   b    : the right-hand side (br = length, bc = 1 for a vector; 0 0 when unused)
//   values may be nan / inf / -inf
// Output per case:
//   OK <count> <arrays...>   each array as "<rows> <cols> <values>" (%.17g)
//   ERR <Kind> <Display text with each newline written as <LF>>
#include "scirs2_linalg.hpp"

#include <cstdio>
#include <iostream>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

using namespace scirs2_linalg;

static const char *kind_name(LinalgError::Kind k)
{
    switch (k) {
    case LinalgError::Kind::ShapeError: return "ShapeError";
    case LinalgError::Kind::InvalidInputError: return "InvalidInputError";
    case LinalgError::Kind::SingularMatrixError: return "SingularMatrixError";
    case LinalgError::Kind::NonPositiveDefiniteError: return "NonPositiveDefiniteError";
    case LinalgError::Kind::ComputationError: return "ComputationError";
    case LinalgError::Kind::ValueError: return "ValueError";
    case LinalgError::Kind::NotImplementedError: return "NotImplementedError";
    default: return "Other";
    }
}

static void print_err(const LinalgError &e)
{
    std::string t = e.to_string(), o;
    for (char ch : t) {
        if (ch == '\n') o += "<LF>";
        else o += ch;
    }
    std::printf("ERR %s %s\n", kind_name(e.kind()), o.c_str());
}

template <class F>
static void put(std::size_t r, std::size_t c, const F *v)
{
    std::printf(" %zu %zu", r, c);
    for (std::size_t i = 0; i < r * c; ++i) std::printf(" %.17g", (double)v[i]);
}

static double read_val(std::istream &in)
{
    std::string s;
    in >> s;
    return std::stod(s);   // accepts nan, inf, -inf
}

template <class F>
static void run(const std::string &op, int f1, int f2, std::size_t ar, std::size_t ac, const std::vector<double> &av,
                std::size_t br, std::size_t bc, const std::vector<double> &bv)
{
    Array2<F> a(ar, ac), b(br, bc);
    for (std::size_t i = 0; i < av.size(); ++i) a.data[i] = (F)av[i];
    for (std::size_t i = 0; i < bv.size(); ++i) b.data[i] = (F)bv[i];
    std::vector<F> bvec(b.data.begin(), b.data.end());
    const std::optional<std::size_t> workers;
    if (op == "solve") {
        auto r = solve<F>(a.view(), view1(bvec), workers);
        if (r.is_err()) return print_err(r.error());
        std::printf("OK 1");
        put(r.value().size(), 1, r.value().data());
    } else if (op == "solve_triangular") {
        auto r = solve_triangular<F>(a.view(), view1(bvec), f1 != 0, f2 != 0);
        if (r.is_err()) return print_err(r.error());
        std::printf("OK 1");
        put(r.value().size(), 1, r.value().data());
    } else if (op == "solve_multiple") {
        auto r = solve_multiple<F>(a.view(), b.view(), workers);
        if (r.is_err()) return print_err(r.error());
        std::printf("OK 1");
        put(r.value().rows, r.value().cols, r.value().data.data());
    } else if (op == "lu") {
        auto r = lu<F>(a.view(), workers);
        if (r.is_err()) return print_err(r.error());
        const auto &[p, l, u] = r.value();
        std::printf("OK 3");
        put(p.rows, p.cols, p.data.data());
        put(l.rows, l.cols, l.data.data());
        put(u.rows, u.cols, u.data.data());
    } else if (op == "cholesky") {
        auto r = cholesky<F>(a.view(), workers);
        if (r.is_err()) return print_err(r.error());
        std::printf("OK 1");
        put(r.value().rows, r.value().cols, r.value().data.data());
    } else if (op == "det") {
        auto r = det<F>(a.view(), workers);
        if (r.is_err()) return print_err(r.error());
        F v = r.value();
        std::printf("OK 1");
        put(1, 1, &v);
    } else if (op == "inv") {
        auto r = inv<F>(a.view(), workers);
        if (r.is_err()) return print_err(r.error());
        std::printf("OK 1");
        put(r.value().rows, r.value().cols, r.value().data.data());
    } else if (op == "mixed_precision_solve") {
        // f1 selects the elimination precision H: 0 f64, 1 f32 (A, B, C = the case type)
        auto r = f1 ? mixed_precision_solve<F, F, F, float>(a.view(), view1(bvec))
                    : mixed_precision_solve<F, F, F, double>(a.view(), view1(bvec));
        if (r.is_err()) return print_err(r.error());
        std::printf("OK 1");
        put(r.value().size(), 1, r.value().data());
    } else if (op == "iterative_refinement_solve") {
        // H = f64; f1 selects the working precision W: 0 f32, 1 f64; f2 = max_iter (0: default)
        const std::optional<std::size_t> it = f2 > 0 ? std::optional<std::size_t>(f2) : std::nullopt;
        auto r = f1 ? iterative_refinement_solve<F, F, F, double, double>(a.view(), view1(bvec), it, std::nullopt)
                    : iterative_refinement_solve<F, F, F, double, float>(a.view(), view1(bvec), it, std::nullopt);
        if (r.is_err()) return print_err(r.error());
        std::printf("OK 1");
        put(r.value().size(), 1, r.value().data());
    } else if (op == "compat_solve") {
        // f1 bits: 1 lower, 2 transposed, 4 check_finite; f2: 0 no assume_a, 1 "pos", 2 "sym"
        std::optional<std::string> assume;
        if (f2 == 1) assume = "pos";
        else if (f2 == 2) assume = "sym";
        auto r = compat_solve<F>(a.view(), b.view(), (f1 & 1) != 0, false, false, (f1 & 4) != 0, assume, (f1 & 2) != 0);
        if (r.is_err()) return print_err(r.error());
        std::printf("OK 1");
        put(r.value().rows, r.value().cols, r.value().data.data());
    } else if (op == "gmres") {
        // f1 = max_iterations (0: default 1000), f2 = restart (0: default); b = right-hand side
        if constexpr (std::is_same<F, double>::value) {
            IterativeSolverOptions<F> opt;
            if (f1 > 0) opt.max_iterations = (std::size_t)f1;
            if (f2 > 0) opt.restart = (std::size_t)f2;
            else if (f2 < 0) opt.restart = (std::size_t)0;   // an explicit restart of 0
            auto r = gmres<F>(a.view(), view1(bvec), nullptr, opt);
            if (r.is_err()) return print_err(r.error());
            const auto &v = r.value();
            const double itv = (double)v.iterations, cv = v.converged ? 1.0 : 0.0;
            std::printf("OK 4");
            put(v.solution.size(), 1, v.solution.data());
            put(1, 1, &itv);
            put(1, 1, &v.residual_norm);
            put(1, 1, &cv);
        } else {
            std::printf("ERR Other gmres is f64 only\n");
            return;
        }
    } else if (op == "preconditioned_gmres") {
        // f1 = preconditioner: 0 Diagonal, 1 LU; f2 = restart (< 0: 0); max_iterations 1000, tol 1e-10
        if constexpr (std::is_same<F, double>::value) {
            const std::size_t restart = f2 < 0 ? 0 : (f2 > 0 ? (std::size_t)f2 : 30);
            Result<Array1<F>> r = [&]() -> Result<Array1<F>> {
                if (f1 == 1) {
                    LUPreconditioner<F> lu(a.rows);
                    return preconditioned_gmres<F>(a.view(), view1(bvec), lu, 1000, 1e-10, restart, nullptr);
                }
                auto dp = DiagonalPreconditioner<F>::create(a.view());
                if (dp.is_err()) return dp.error();
                return preconditioned_gmres<F>(a.view(), view1(bvec), dp.value(), 1000, 1e-10, restart, nullptr);
            }();
            if (r.is_err()) return print_err(r.error());
            std::printf("OK 1");
            put(r.value().size(), 1, r.value().data());
        } else {
            std::printf("ERR Other preconditioned_gmres is f64 only\n");
            return;
        }
    } else if (op == "batch_matmul" || op == "batch_matvec" || op == "batch_add" || op == "batch_sum") {
        // f1 = batch count: A is given as (batch*m) x k; f2 = axis for batch_add
        const std::size_t nb = (std::size_t)f1, m = nb ? ar / nb : 0;
        ArrayView3<F> a3{a.data.data(), nb, m, ac, (std::ptrdiff_t)(m * ac), (std::ptrdiff_t)ac, 1};
        if (op == "batch_matmul") {
            auto r = batch_matmul<F>(a3, b.view());
            if (r.is_err()) return print_err(r.error());
            std::printf("OK 1");
            put(r.value().d0 * r.value().d1, r.value().d2, r.value().data.data());
        } else if (op == "batch_matvec") {
            auto r = batch_matvec<F>(a3, view1(bvec));
            if (r.is_err()) return print_err(r.error());
            std::printf("OK 1");
            put(r.value().rows, r.value().cols, r.value().data.data());
        } else if (op == "batch_add") {
            auto r = batch_add<F>(a3, view1(bvec), (std::size_t)f2);
            if (r.is_err()) return print_err(r.error());
            std::printf("OK 1");
            put(r.value().d0 * r.value().d1, r.value().d2, r.value().data.data());
        } else {
            auto r = batch_sum<F>(a3);
            if (r.is_err()) return print_err(r.error());
            std::printf("OK 1");
            put(r.value().rows, r.value().cols, r.value().data.data());
        }
    } else if (op == "qr" || op == "lstsq") {
        if constexpr (std::is_same<F, double>::value) {
            if (op == "qr") {
                auto r = qr<F>(a.view(), workers);
                if (r.is_err()) return print_err(r.error());
                std::printf("OK 2");
                put(r.value().first.rows, r.value().first.cols, r.value().first.data.data());
                put(r.value().second.rows, r.value().second.cols, r.value().second.data.data());
            } else {
                auto r = lstsq<F>(a.view(), view1(bvec), workers);
                if (r.is_err()) return print_err(r.error());
                const auto &v = r.value();
                const double rk = (double)v.rank;
                std::printf("OK 4");
                put(v.x.size(), 1, v.x.data());
                put(1, 1, &v.residuals);
                put(1, 1, &rk);
                put(v.s.size(), 1, v.s.data());
            }
        } else {
            std::printf("ERR Other %s is f64 only\n", op.c_str());
            return;
        }
    } else {
        std::printf("ERR Other unknown op %s\n", op.c_str());
        return;
    }
    std::printf("\n");
}

int main()
{
    std::string op, type;
    while (std::cin >> op >> type) {
        int f1, f2;
        std::size_t ar, ac, br, bc;
        std::cin >> f1 >> f2 >> ar >> ac;
        std::vector<double> av(ar * ac);
        for (auto &v : av) v = read_val(std::cin);
        std::cin >> br >> bc;
        std::vector<double> bv(br * bc);
        for (auto &v : bv) v = read_val(std::cin);
        if (type == "f32") run<float>(op, f1, f2, ar, ac, av, br, bc, bv);
        else run<double>(op, f1, f2, ar, ac, av, br, bc, bv);
        std::fflush(stdout);
    }
    return 0;
}
