// TEST INFRASTRUCTURE ONLY: command-line driver around the reference's own
// FP64 dual simplex (core/src/simplex, built from the reference sources by
// oracle/ref/Makefile into oracle/_ref/).  Used by tests/golden/make_fixtures.py
// to write objective / primal / dual / basis fixtures; never shipped.
//
// Input (text, whitespace separated; "inf"/"-inf" allowed):
//   m n nnz | col_ptr[n+1] | row_ind[nnz] | vals[nnz] | c[n] | l[n] | u[n] | rl[m] | ru[m]
// Or: refsimplex --mps file.mps dump.txt -- parse an MPS file with the
// reference's own reader (core/src/io/mps_io.cpp), write it in the format
// above (plus a first line "sense S") and solve it.
// Output: "status S", "objective V", "iterations K", then x, row_duals,
// reduced_costs, col_basis, row_basis, one vector per line.
// Or: refsimplex --mps-dump file.mps dump.txt -- parse only, and write every
// field of the parsed model::Problem (name, sense, offset, CSC, bounds,
// variable types, names, quadratic CSC); on a parse error print
// "error <what()>" and exit 3.
// Or: refsimplex --mps-write in.mps out.mps -- parse with the reference reader
// and write the problem back with the reference writer (io::write_mps).
// Or: refsimplex --from-basis lp.txt basis.txt -- the LP as above, and a basis
// (col_basis[n] then row_basis[m], BasisStatus integers); runs
// SimplexSolver::solve_from_basis with the options GpuSolver::solve_hybrid
// passes (gpu_solver.cpp:397-402) and prints the solution as above.
// Or: refsimplex --check lp.txt sol.txt -- the LP as above followed by
// "sense S offset O", "vtypes" and n integers, "q NNZ" with Q's CSC arrays
// (NNZ 0: no Q); the solution file holds "status S", "objective V", then
// "x", "row_duals", "reduced_costs", "ray" each as a count and values, and
// "tols tp td ti".  Runs checker::check_solution and prints every
// CheckResult field (doubles in %a) and the summary.
#include <sih/checker/checker.hpp>
#include <sih/io/mps_io.hpp>
#include <sih/model/problem.hpp>
#include <sih/model/options.hpp>
#include <sih/simplex/simplex_solver.hpp>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

static double rd(std::istream &in) {
    std::string t;
    in >> t;
    if (t == "inf") return std::numeric_limits<double>::infinity();
    if (t == "-inf") return -std::numeric_limits<double>::infinity();
    return std::stod(t);
}

static void dump(const sih::model::Problem &p, const char *path) {
    std::FILE *f = std::fopen(path, "w");
    const auto &A = p.A();
    std::fprintf(f, "sense %d\n%lld %lld %lld\n", static_cast<int>(p.sense()),
                 static_cast<long long>(p.num_rows()), static_cast<long long>(p.num_cols()),
                 static_cast<long long>(p.num_nonzeros()));
    auto ivec = [&](const std::vector<int64_t> &x) { for (auto e : x) std::fprintf(f, "%lld ", static_cast<long long>(e)); std::fprintf(f, "\n"); };
    auto dvec = [&](const std::vector<double> &x) {
        for (double e : x) {
            if (std::isinf(e)) std::fprintf(f, "%s ", e > 0 ? "inf" : "-inf");
            else std::fprintf(f, "%.17g ", e);
        }
        std::fprintf(f, "\n");
    };
    ivec(A.csc_col_ptr()); ivec(A.csc_row_ind()); dvec(A.csc_values());
    dvec(p.c()); dvec(p.col_lower()); dvec(p.col_upper()); dvec(p.row_lower()); dvec(p.row_upper());
    std::fclose(f);
}

static void full_dump(const sih::model::Problem &p, const char *path) {
    std::FILE *f = std::fopen(path, "w");
    auto ivec = [&](const char *k, const std::vector<int64_t> &x) {
        std::fprintf(f, "%s %zu", k, x.size());
        for (auto e : x) std::fprintf(f, " %lld", static_cast<long long>(e));
        std::fprintf(f, "\n");
    };
    auto dvec = [&](const char *k, const std::vector<double> &x) {
        std::fprintf(f, "%s %zu", k, x.size());
        for (double e : x) {
            if (std::isinf(e)) std::fprintf(f, " %s", e > 0 ? "inf" : "-inf");
            else if (std::isnan(e)) std::fprintf(f, " nan");
            else std::fprintf(f, " %a", e);
        }
        std::fprintf(f, "\n");
    };
    auto names = [&](const char *k, const std::vector<std::string> &x) {
        std::fprintf(f, "%s %zu\n", k, x.size());
        for (const auto &e : x) std::fprintf(f, "%s\n", e.c_str());
    };
    std::fprintf(f, "name\n%s\n", p.name().c_str());
    std::fprintf(f, "sense %d\nobj_offset %a\n", static_cast<int>(p.sense()), p.obj_offset());
    std::fprintf(f, "dims %lld %lld %lld\n", static_cast<long long>(p.num_rows()),
                 static_cast<long long>(p.num_cols()), static_cast<long long>(p.num_nonzeros()));
    ivec("col_ptr", p.A().csc_col_ptr());
    ivec("row_ind", p.A().csc_row_ind());
    dvec("vals", p.A().csc_values());
    dvec("c", p.c());
    dvec("col_lower", p.col_lower());
    dvec("col_upper", p.col_upper());
    dvec("row_lower", p.row_lower());
    dvec("row_upper", p.row_upper());
    std::vector<int64_t> vt;
    for (auto t : p.var_types()) vt.push_back(static_cast<int64_t>(t));
    ivec("var_types", vt);
    names("row_names", p.row_names());
    names("col_names", p.col_names());
    std::fprintf(f, "has_quadratic %d\n", p.has_quadratic() ? 1 : 0);
    if (p.has_quadratic()) {
        ivec("q_col_ptr", p.Q().csc_col_ptr());
        ivec("q_row_ind", p.Q().csc_row_ind());
        dvec("q_vals", p.Q().csc_values());
    }
    std::fclose(f);
}

static void solve_and_print(const sih::model::Problem &p);
static void print_solution(const sih::model::Solution &s);

int main(int argc, char **argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: refsimplex lp.txt | --mps file.mps dump.txt\n"); return 2; }
    if (std::string(argv[1]) == "--mps-dump" && argc >= 4) {
        try {
            full_dump(sih::io::read_mps(argv[2]), argv[3]);
        } catch (const std::exception &e) {
            std::printf("error %s\n", e.what());
            return 3;
        }
        return 0;
    }
    if (std::string(argv[1]) == "--mps-write" && argc >= 4) {
        try {
            sih::io::write_mps(sih::io::read_mps(argv[2]), argv[3]);
        } catch (const std::exception &e) {
            std::printf("error %s\n", e.what());
            return 3;
        }
        return 0;
    }
    if (std::string(argv[1]) == "--mps" && argc >= 4) {
        sih::model::Problem p = sih::io::read_mps(argv[2]);
        dump(p, argv[3]);
        solve_and_print(p);
        return 0;
    }
    const bool from_basis = std::string(argv[1]) == "--from-basis" && argc >= 4;
    const bool check = std::string(argv[1]) == "--check" && argc >= 4;
    std::ifstream in(argv[(from_basis || check) ? 2 : 1]);
    long m, n, nnz;
    in >> m >> n >> nnz;
    std::vector<long> cp(n + 1), ri(nnz);
    std::vector<double> v(nnz);
    for (auto &e : cp) in >> e;
    for (auto &e : ri) in >> e;
    for (auto &e : v) e = rd(in);
    sih::model::Problem p("fixture");
    p.resize(m, n);
    std::vector<sih::model::Triplet> trip;
    for (long j = 0; j < n; ++j)
        for (long k = cp[j]; k < cp[j + 1]; ++k) trip.emplace_back(ri[k], j, v[k]);
    p.set_A(sih::model::SparseMatrix::from_triplets(m, n, trip));
    std::vector<double> c(n);
    for (auto &e : c) e = rd(in);
    p.set_c(c);
    for (long j = 0; j < n; ++j) p.col_lower()[j] = rd(in);
    for (long j = 0; j < n; ++j) p.col_upper()[j] = rd(in);
    for (long i = 0; i < m; ++i) p.row_lower()[i] = rd(in);
    for (long i = 0; i < m; ++i) p.row_upper()[i] = rd(in);
    if (check) {
        std::string tag;
        int sense;
        in >> tag >> sense >> tag;
        const double offset = rd(in);
        p.set_sense(static_cast<sih::model::ObjectiveSense>(sense));
        p.set_obj_offset(offset);
        in >> tag;
        for (long j = 0; j < n; ++j) { int t; in >> t; p.var_types()[j] = static_cast<sih::model::VariableType>(t); }
        long qnz;
        in >> tag >> qnz;
        if (qnz > 0) {
            std::vector<long> qp(n + 1), qi(qnz);
            std::vector<double> qv(qnz);
            for (auto &e : qp) in >> e;
            for (auto &e : qi) in >> e;
            for (auto &e : qv) e = rd(in);
            std::vector<sih::model::Triplet> qt;
            for (long j = 0; j < n; ++j)
                for (long k = qp[j]; k < qp[j + 1]; ++k) qt.emplace_back(qi[k], j, qv[k]);
            p.set_quadratic_objective(sih::model::SparseMatrix::from_triplets(n, n, qt));
        }
        std::ifstream sin(argv[3]);
        sih::model::Solution sol;
        int st;
        sin >> tag >> st >> tag;
        sol.status = static_cast<sih::model::SolutionStatus>(st);
        sol.primal_objective = rd(sin);
        auto vec = [&](std::vector<double> &x) {
            long k;
            sin >> tag >> k;
            x.resize(static_cast<size_t>(k));
            for (auto &e : x) e = rd(sin);
        };
        vec(sol.x);
        vec(sol.row_duals);
        vec(sol.reduced_costs);
        vec(sol.ray);
        double tp, td, ti;
        sin >> tag;
        tp = rd(sin);
        td = rd(sin);
        ti = rd(sin);
        const sih::checker::CheckResult r = sih::checker::check_solution(p, sol, tp, td, ti);
        std::printf("is_primal_feasible %d\nmax_row_violation %a\nmax_col_bound_violation %a\n"
                    "max_primal_residual %a\nis_dual_feasible %d\nmax_dual_residual %a\n"
                    "is_complementary %d\nmax_complementarity_slack %a\nis_integrality_satisfied %d\n"
                    "max_integrality_violation %a\nevaluated_objective %a\nobjective_discrepancy %a\n"
                    "is_certificate_valid %d\ncertificate_violation %a\nall_checks_passed %d\nsummary %s\n",
                    r.is_primal_feasible, r.max_row_violation, r.max_col_bound_violation, r.max_primal_residual,
                    r.is_dual_feasible, r.max_dual_residual, r.is_complementary, r.max_complementarity_slack,
                    r.is_integrality_satisfied, r.max_integrality_violation, r.evaluated_objective,
                    r.objective_discrepancy, r.is_certificate_valid, r.certificate_violation,
                    r.all_checks_passed, r.summary.c_str());
        return 0;
    }
    if (from_basis) {
        std::ifstream bin(argv[3]);
        std::vector<sih::model::BasisStatus> cb(n), rb(m);
        for (auto &e : cb) { int v; bin >> v; e = static_cast<sih::model::BasisStatus>(v); }
        for (auto &e : rb) { int v; bin >> v; e = static_cast<sih::model::BasisStatus>(v); }
        sih::model::Options opts;
        opts.strategy.algorithm = sih::model::AlgorithmChoice::DualSimplex;
        opts.strategy.presolve = sih::model::PresolveMode::Off;
        opts.strategy.enable_scaling = false;
        opts.log_to_console = false;
        opts.telemetry_log_path = "";
        print_solution(sih::simplex::SimplexSolver::solve_from_basis(p, cb, rb, opts));
        return 0;
    }
    solve_and_print(p);
    return 0;
}

static void solve_and_print(const sih::model::Problem &p) {
    sih::model::Options opts;
    opts.log_to_console = false;
    opts.telemetry_log_path = "";
    opts.strategy.algorithm = sih::model::AlgorithmChoice::DualSimplex;
    print_solution(sih::simplex::SimplexSolver::solve(p, opts));
}

static void print_solution(const sih::model::Solution &s) {
    std::printf("status %d\nobjective %.17g\niterations %lld\n", static_cast<int>(s.status),
                s.primal_objective, static_cast<long long>(s.simplex_iterations));
    auto vec = [](const char *name, const std::vector<double> &x) {
        std::printf("%s", name);
        for (double e : x) std::printf(" %.17g", e);
        std::printf("\n");
    };
    vec("x", s.x);
    vec("row_duals", s.row_duals);
    vec("reduced_costs", s.reduced_costs);
    std::printf("col_basis");
    for (auto b : s.col_basis) std::printf(" %d", static_cast<int>(b));
    std::printf("\nrow_basis");
    for (auto b : s.row_basis) std::printf(" %d", static_cast<int>(b));
    std::printf("\n");
}
