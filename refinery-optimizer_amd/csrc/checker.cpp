// The solution checker of the reference (checker::check_solution,
// core/src/checker/checker.cpp:9-268) over the C ABI's LP and solution
// types: the same checks, tolerances and result fields, with its arithmetic
// order -- A x as row sums in CSR order and A^T y as column sums in CSC order
// (SparseMatrix::mat_vec / mat_trans_vec, sparse_matrix.cpp:90-129), every
// other sum sequential -- so that each figure equals the reference's.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "mpdhg.h"
#include "mpdhg_internal.hpp"

namespace mpdhg {
namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

// CSR view (row pointers, columns ascending within a row, values) of a CSC.
struct Csr {
    std::vector<int64_t> ptr, col;
    std::vector<double> val;
};

Csr to_csr(int64_t m, int64_t n, const int64_t *cp, const int64_t *ri, const double *v) {
    Csr r;
    r.ptr.assign(static_cast<size_t>(m) + 1, 0);
    const int64_t nnz = n > 0 ? cp[n] : 0;
    for (int64_t k = 0; k < nnz; ++k) r.ptr[static_cast<size_t>(ri[k]) + 1]++;
    for (int64_t i = 0; i < m; ++i) r.ptr[static_cast<size_t>(i) + 1] += r.ptr[static_cast<size_t>(i)];
    r.col.resize(static_cast<size_t>(nnz));
    r.val.resize(static_cast<size_t>(nnz));
    std::vector<int64_t> fill(r.ptr.begin(), r.ptr.end() - 1);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = cp[j]; k < cp[j + 1]; ++k) {
            const int64_t d = fill[static_cast<size_t>(ri[k])]++;
            r.col[static_cast<size_t>(d)] = j;
            r.val[static_cast<size_t>(d)] = v[k];
        }
    return r;
}

std::vector<double> mat_vec(const Csr &a, int64_t m, const double *x) {
    std::vector<double> y(static_cast<size_t>(m), 0.0);
    for (int64_t i = 0; i < m; ++i) {
        double sum = 0.0;
        for (int64_t k = a.ptr[static_cast<size_t>(i)]; k < a.ptr[static_cast<size_t>(i) + 1]; ++k)
            sum += a.val[static_cast<size_t>(k)] * x[a.col[static_cast<size_t>(k)]];
        y[static_cast<size_t>(i)] = sum;
    }
    return y;
}

std::vector<double> mat_trans_vec(const mpdhg_lp &lp, const double *y) {
    std::vector<double> z(static_cast<size_t>(lp.n), 0.0);
    for (int64_t j = 0; j < lp.n; ++j) {
        double sum = 0.0;
        for (int64_t k = lp.col_ptr[j]; k < lp.col_ptr[j + 1]; ++k) sum += lp.vals[k] * y[lp.row_ind[k]];
        z[static_cast<size_t>(j)] = sum;
    }
    return z;
}

// std::ostream's default formatting of a double (6 significant digits).
std::string num(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%g", v);
    return b;
}

void set_summary(mpdhg_check_result &r, const std::string &s) {
    std::snprintf(r.summary, sizeof(r.summary), "%s", s.c_str());
}

void check(const mpdhg_lp &lp, const mpdhg_lp_meta *meta, const mpdhg_check_input &sol, double tol_primal,
           double tol_dual, double tol_int, mpdhg_check_result &res) {
    const int64_t m = lp.m, n = lp.n;
    const int32_t sense = meta ? meta->sense : 1;
    const double offset = meta ? meta->obj_offset : 0.0;
    const bool has_q = meta && meta->has_quadratic && meta->q_nnz > 0;
    const Csr A = to_csr(m, n, lp.col_ptr, lp.row_ind, lp.vals);

    // 0. certificates (:18-139)
    if (sol.status == MPDHG_STATUS_INFEASIBLE) {
        if (!sol.ray || sol.ray_len == 0 || sol.ray_len != m) {
            set_summary(res, "Infeasible certificate error: Farkas ray dimension mismatch or empty");
            res.all_checks_passed = 0;
            return;
        }
        double ray_norm = 0.0;
        for (int64_t i = 0; i < m; ++i) ray_norm = std::max(ray_norm, std::abs(sol.ray[i]));
        if (ray_norm < 1e-12) {
            set_summary(res, "Infeasible certificate error: Farkas ray is zero");
            res.all_checks_passed = 0;
            return;
        }
        bool valid = false;
        double best_gap = -1.0;
        for (double sign : {1.0, -1.0}) {
            std::vector<double> y(static_cast<size_t>(m));
            for (int64_t i = 0; i < m; ++i) y[static_cast<size_t>(i)] = sol.ray[i] * sign;
            const std::vector<double> w = mat_trans_vec(lp, y.data());
            bool inf_lhs = false;
            double lhs = 0.0;
            for (int64_t i = 0; i < m; ++i) {
                const double yi = y[static_cast<size_t>(i)];
                if (yi > 1e-9) {
                    if (lp.row_lower[i] <= -kInf / 2.0) { inf_lhs = true; break; }
                    lhs += yi * lp.row_lower[i];
                } else if (yi < -1e-9) {
                    if (lp.row_upper[i] >= kInf / 2.0) { inf_lhs = true; break; }
                    lhs += yi * lp.row_upper[i];
                }
            }
            if (inf_lhs) continue;
            bool inf_rhs = false;
            double rhs = 0.0;
            for (int64_t j = 0; j < n; ++j) {
                const double wj = w[static_cast<size_t>(j)];
                if (wj > 1e-9) {
                    if (lp.col_upper[j] >= kInf / 2.0) { inf_rhs = true; break; }
                    rhs += wj * lp.col_upper[j];
                } else if (wj < -1e-9) {
                    if (lp.col_lower[j] <= -kInf / 2.0) { inf_rhs = true; break; }
                    rhs += wj * lp.col_lower[j];
                }
            }
            if (inf_rhs) continue;
            const double gap = lhs - rhs;
            if (gap > 1e-6) {
                valid = true;
                best_gap = std::max(best_gap, gap);
            }
        }
        res.is_certificate_valid = valid;
        res.certificate_violation = valid ? 0.0 : 1.0;
        res.all_checks_passed = valid;
        set_summary(res, valid ? "Checker Summary: PASSED | Farkas certificate ray verified"
                               : "Checker Summary: FAILED | Farkas certificate ray invalid");
        return;
    }
    if (sol.status == MPDHG_STATUS_UNBOUNDED) {
        if (!sol.ray || sol.ray_len == 0 || sol.ray_len != n) {
            set_summary(res, "Unbounded certificate error: ray dimension mismatch or empty");
            res.all_checks_passed = 0;
            return;
        }
        double ray_norm = 0.0;
        for (int64_t j = 0; j < n; ++j) ray_norm = std::max(ray_norm, std::abs(sol.ray[j]));
        if (ray_norm < 1e-12) {
            set_summary(res, "Unbounded certificate error: ray is zero");
            res.all_checks_passed = 0;
            return;
        }
        double c_dot_d = 0.0;
        for (int64_t j = 0; j < n; ++j) c_dot_d += lp.c[j] * sol.ray[j];
        const bool improves = (sense == -1) ? (c_dot_d > 1e-9) : (c_dot_d < -1e-9);
        bool bounds_ok = true;
        for (int64_t j = 0; j < n; ++j) {
            if (sol.ray[j] > 1e-9 && lp.col_upper[j] < kInf / 2.0) { bounds_ok = false; break; }
            if (sol.ray[j] < -1e-9 && lp.col_lower[j] > -kInf / 2.0) { bounds_ok = false; break; }
        }
        const std::vector<double> ad = mat_vec(A, m, sol.ray);
        bool rows_ok = true;
        for (int64_t i = 0; i < m; ++i) {
            const double v = ad[static_cast<size_t>(i)];
            if (v > 1e-9 && lp.row_upper[i] < kInf / 2.0) { rows_ok = false; break; }
            if (v < -1e-9 && lp.row_lower[i] > -kInf / 2.0) { rows_ok = false; break; }
        }
        const bool valid = improves && bounds_ok && rows_ok;
        res.is_certificate_valid = valid;
        res.certificate_violation = valid ? 0.0 : 1.0;
        res.all_checks_passed = valid;
        set_summary(res, valid ? "Checker Summary: PASSED | Unbounded ray certificate verified"
                               : "Checker Summary: FAILED | Unbounded ray certificate invalid");
        return;
    }
    const int64_t x_len = sol.x ? n : 0;
    if (x_len != n) {
        set_summary(res, "Primal vector x dimension mismatch");
        return;
    }

    // 1. primal feasibility (:146-183)
    const std::vector<double> ax = mat_vec(A, m, sol.x);
    double max_row = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        const double lb = lp.row_lower[i], ub = lp.row_upper[i], v = ax[static_cast<size_t>(i)];
        if (lb > -kInf / 2.0 && v < lb) max_row = std::max(max_row, lb - v);
        if (ub < kInf / 2.0 && v > ub) max_row = std::max(max_row, v - ub);
    }
    res.max_row_violation = max_row;
    double max_col = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        const double lb = lp.col_lower[j], ub = lp.col_upper[j], v = sol.x[j];
        if (lb > -kInf / 2.0 && v < lb) max_col = std::max(max_col, lb - v);
        if (ub < kInf / 2.0 && v > ub) max_col = std::max(max_col, v - ub);
    }
    res.max_col_bound_violation = max_col;
    res.max_primal_residual = std::max(max_row, max_col);
    res.is_primal_feasible = res.max_primal_residual <= tol_primal;

    // 2. integrality (:185-196)
    bool is_mip = false;
    double max_int = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        const int8_t vt = (meta && meta->var_type) ? meta->var_type[j] : MPDHG_VAR_CONTINUOUS;
        if (vt == MPDHG_VAR_INTEGER || vt == MPDHG_VAR_BINARY) {
            is_mip = true;
            max_int = std::max(max_int, std::abs(sol.x[j] - std::round(sol.x[j])));
        }
    }
    res.max_integrality_violation = max_int;
    res.is_integrality_satisfied = max_int <= tol_int;

    // 3. objective (:198-213)
    Csr Q;
    std::vector<double> qx;
    if (has_q) {
        Q = to_csr(n, n, meta->q_col_ptr, meta->q_row_ind, meta->q_vals);
        qx = mat_vec(Q, n, sol.x);
    }
    double obj = offset;
    for (int64_t j = 0; j < n; ++j) obj += lp.c[j] * sol.x[j];
    if (has_q) {
        double quad = 0.0;
        for (int64_t j = 0; j < n; ++j) quad += sol.x[j] * qx[static_cast<size_t>(j)];
        obj += 0.5 * quad;
    }
    res.evaluated_objective = obj;
    res.objective_discrepancy = std::abs(obj - sol.primal_objective);

    // 4. stationarity and complementarity (:215-252)
    const bool has_duals = sol.row_duals || m == 0;
    const bool has_rc = sol.reduced_costs || n == 0;
    if (!is_mip && has_duals && has_rc) {
        const std::vector<double> aty = mat_trans_vec(lp, sol.row_duals);
        double max_d = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            double grad = lp.c[j];
            if (has_q) grad += qx[static_cast<size_t>(j)];
            max_d = std::max(max_d, std::abs(grad - aty[static_cast<size_t>(j)] - sol.reduced_costs[j]));
        }
        res.max_dual_residual = max_d;
        res.is_dual_feasible = max_d <= tol_dual;
        double max_cs = 0.0;
        for (int64_t i = 0; i < m; ++i) {
            const double act = ax[static_cast<size_t>(i)];
            const double dist = std::min(std::abs(act - lp.row_lower[i]), std::abs(lp.row_upper[i] - act));
            max_cs = std::max(max_cs, dist * std::abs(sol.row_duals[i]));
        }
        res.max_complementarity_slack = max_cs;
        res.is_complementary = max_cs <= tol_dual * 10.0;
    } else {
        res.is_dual_feasible = 1;
        res.is_complementary = 1;
    }

    const bool obj_match = res.objective_discrepancy <= 1e-4 * (1.0 + std::abs(obj));
    res.all_checks_passed = res.is_primal_feasible && res.is_integrality_satisfied && obj_match;
    set_summary(res, std::string("Checker Summary: ") + (res.all_checks_passed ? "PASSED" : "FAILED") +
                         " | Max Primal Viol: " + num(res.max_primal_residual) +
                         " | Max Int Viol: " + num(res.max_integrality_violation) +
                         " | Obj Evaluated: " + num(res.evaluated_objective) +
                         " | Obj Discrepancy: " + num(res.objective_discrepancy));
}

}  // namespace
}  // namespace mpdhg

extern "C" int mpdhg_check_solution(const mpdhg_lp *lp, const mpdhg_lp_meta *meta, const mpdhg_check_input *sol,
                                    double tol_primal, double tol_dual, double tol_int,
                                    mpdhg_check_result *out) {
    if (!lp || !sol || !out) return mpdhg::fail(MPDHG_E_ARG, "check_solution: NULL argument");
    if (lp->m < 0 || lp->n < 0 || lp->nnz < 0 || (lp->n > 0 && !lp->col_ptr) ||
        (lp->nnz > 0 && (!lp->row_ind || !lp->vals)))
        return mpdhg::fail(MPDHG_E_ARG, "check_solution: malformed LP");
    if (lp->n > 0 && (lp->col_ptr[0] != 0 || lp->col_ptr[lp->n] != lp->nnz))
        return mpdhg::fail(MPDHG_E_ARG, "check_solution: col_ptr must start at 0 and end at nnz");
    for (int64_t j = 0; j < lp->n; ++j)
        if (lp->col_ptr[j + 1] < lp->col_ptr[j]) return mpdhg::fail(MPDHG_E_ARG, "check_solution: col_ptr not monotone");
    for (int64_t k = 0; k < lp->nnz; ++k)
        if (lp->row_ind[k] < 0 || lp->row_ind[k] >= lp->m)
            return mpdhg::fail(MPDHG_E_ARG, "check_solution: row index out of range");
    if (meta && meta->has_quadratic && meta->q_nnz > 0) {
        if (!meta->q_col_ptr || !meta->q_row_ind || !meta->q_vals || meta->q_col_ptr[0] != 0 ||
            meta->q_col_ptr[lp->n] != meta->q_nnz)
            return mpdhg::fail(MPDHG_E_ARG, "check_solution: malformed Q");
        for (int64_t k = 0; k < meta->q_nnz; ++k)
            if (meta->q_row_ind[k] < 0 || meta->q_row_ind[k] >= lp->n)
                return mpdhg::fail(MPDHG_E_ARG, "check_solution: Q index out of range");
    }
    std::memset(out, 0, sizeof(*out));
    try {
        mpdhg::check(*lp, meta, *sol, tol_primal, tol_dual, tol_int, *out);
    } catch (const std::bad_alloc &) {
        return mpdhg::fail(MPDHG_E_NOMEM, "check_solution: out of host memory");
    }
    return MPDHG_OK;
}
