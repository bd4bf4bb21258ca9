// FP64 crossover: a bounded primal simplex warm-started from a basis, on the
// host.  This is the polish stage of GpuSolver::solve_hybrid
// (gpu_solver.cpp:397-417), which the reference runs as
// SimplexSolver::solve_from_basis (simplex_solver.cpp:266-329) on the CPU after
// the GPU PDHG.  The PDHG itself stays on the device; this stage is sequential
// pivoting on a sparse basis factorisation, which is host work.
//
// Model (the reference's, primal_simplex.cpp:10-47): variables z = (x, r),
// x in [l, u] (n columns), r in [rl, ru] (m row activities), A x - r = 0, so
// the column of r_i is -e_i.  Minimise c^T x.  Duals y = B^-T c_B; reduced
// costs d_j = c_j - a_j^T y, and d = y_i for the row variable r_i.  Bounds
// beyond +-1e20 are infinite (is_bounded_*, problem.hpp:27-35).
//
// Method (own design):
//   * basis factorisation: right-looking sparse LU with Markowitz pivot
//     choice and threshold partial pivoting (|a| >= 0.1 max of its column),
//     count-bucketed search; singletons cost nothing, so the slack and
//     triangular parts of a basis factor without fill.  A column whose
//     remaining entries fall below 1e-7 of its largest original entry is
//     taken as dependent;
//   * updates: product-form eta file, refactorised every `refactor_every`
//     pivots;
//   * pricing: Devex reference weights, largest d_j^2 / w_j (the reference
//     prices by the largest |d_j|);
//   * phase 1: composite, on the sum of the basic variables' bound
//     violations (the cost is -1 below, +1 above, recomputed each
//     iteration); the ratio test stops at the first breakpoint;
//   * ratio test: two-pass (Harris) with the primal tolerance, taking the
//     largest |alpha| among the near-ties, and the entering variable's own
//     bound flip;
//   * degeneracy: every non-fixed finite bound is expanded by a small
//     variable-dependent amount during the solve and restored at the end,
//     where optimality is confirmed again;
//   * a singular warm basis is repaired: each dependent column is replaced
//     by the slack of a row the factorisation left unpivoted.  (The
//     reference falls back to the slack basis instead, primal_simplex.cpp:
//     140-144; both reach an optimal vertex, the repair in fewer pivots.)
//   * crossover from a point x0 (an extension; the hybrid passes the PDHG
//     point): the slack basis with every column at x0, a primal push of the
//     columns inside their bounds, then the simplex.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

#include "mpdhg.h"
#include "mpdhg_internal.hpp"

namespace mpdhg {

namespace {

constexpr double kBig = 1e20;         // |bound| >= this is infinite
constexpr double kPivTol = 1e-9;      // smallest |alpha| a ratio test pivots on
constexpr double kLuThreshold = 0.1;  // threshold partial pivoting
constexpr double kLuTiny = 1e-11;     // a pivot below this is a dependent column
constexpr double kLuDep = 1e-7;       // ... and so is one below this times the column's largest entry
constexpr int kSearch = 4;            // Markowitz candidates examined after the first
constexpr double kPerturb = 1e-7;     // relative bound expansion against degeneracy

enum Stat : int8_t { kBasic = 0, kLower = 1, kUpper = 2, kFree = 3, kSuper = 4 };

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A basis the factorisation could not make nonsingular, even after repair.
struct FactorError : std::runtime_error {
    FactorError() : std::runtime_error("basis factorisation failed after repair") {}
};

// Columns of [A  -I]: j < n structural (CSC of A), j >= n the row variable j-n.
struct Cols {
    int64_t m, n;
    const int64_t *ptr, *idx;
    const double *val;
};

// ---- sparse LU with an eta file -------------------------------------------
class Factor {
  public:
    explicit Factor(int m) : m_(m) {}

    // Factorises the basis whose position k holds variable head[k].  Returns
    // the positions that could not be pivoted (dependent columns) and the
    // rows left without a pivot (as many as positions).
    void factorize(const Cols &A, const std::vector<int64_t> &head, std::vector<int> &bad_pos,
                   std::vector<int> &free_rows);
    // B x = b: b indexed by row in, x indexed by basis position out.
    void ftran(std::vector<double> &v) const;
    // B^T y = c: c indexed by basis position in, y indexed by row out.
    void btran(std::vector<double> &v) const;
    // Position r now holds the column whose B^-1 image is alpha (nonzero at
    // the positions nz, ascending).
    void update(int r, const std::vector<double> &alpha, const std::vector<int> &nz);
    int updates() const { return static_cast<int>(eta_r_.size()); }

  private:
    int m_;
    // pivot sequence: row p_k, position q_k, value piv_k; L_k column
    // multipliers (row, l); U_k row entries (position, u)
    std::vector<int> p_, q_;
    std::vector<double> piv_;
    std::vector<int> lbeg_, urow_beg_;
    std::vector<int> lidx_, uidx_;
    std::vector<double> lval_, uval_;
    // U again by columns (the entries of step k's pivot column in the rows of
    // earlier steps), for a column-oriented FTRAN that skips zero unknowns
    std::vector<int> ucol_beg_, ucol_row_;
    std::vector<double> ucol_val_;
    // eta file
    std::vector<int> eta_r_, eta_beg_, eta_idx_;
    std::vector<double> eta_piv_, eta_val_;
    mutable std::vector<double> w_;
};

// Intrusive doubly linked lists of items bucketed by count.
struct Buckets {
    std::vector<int> head, next, prev, cnt;
    void init(int items, int maxc) {
        head.assign(static_cast<size_t>(maxc + 2), -1);
        next.assign(static_cast<size_t>(items), -1);
        prev.assign(static_cast<size_t>(items), -1);
        cnt.assign(static_cast<size_t>(items), 0);
    }
    void insert(int k, int c) {
        cnt[k] = c;
        prev[k] = -1;
        next[k] = head[c];
        if (head[c] >= 0) prev[head[c]] = k;
        head[c] = k;
    }
    void remove(int k) {
        const int c = cnt[k];
        if (prev[k] >= 0) next[prev[k]] = next[k];
        else head[c] = next[k];
        if (next[k] >= 0) prev[next[k]] = prev[k];
        prev[k] = next[k] = -1;
    }
    void move(int k, int c) {
        remove(k);
        insert(k, c);
    }
};

void Factor::factorize(const Cols &A, const std::vector<int64_t> &head, std::vector<int> &bad_pos,
                       std::vector<int> &free_rows) {
    const int m = m_;
    p_.clear(); q_.clear(); piv_.clear();
    lbeg_.assign(1, 0); urow_beg_.assign(1, 0);
    lidx_.clear(); lval_.clear(); uidx_.clear(); uval_.clear();
    eta_r_.clear(); eta_beg_.assign(1, 0); eta_idx_.clear(); eta_piv_.clear(); eta_val_.clear();
    bad_pos.clear();
    free_rows.clear();
    // Slack positions first: a row variable's column is -e_i, a singleton,
    // so it pivots on row i with no multipliers, and the structural entries
    // of row i become its U row.  Only the structural columns on the
    // remaining rows enter the Markowitz phase.
    std::vector<int> srow_step(static_cast<size_t>(m), -1);
    for (int k = 0; k < m; ++k) {
        const int64_t v = head[static_cast<size_t>(k)];
        if (v < A.n) continue;
        const int i = static_cast<int>(v - A.n);
        srow_step[static_cast<size_t>(i)] = static_cast<int>(p_.size());
        p_.push_back(i);
        q_.push_back(k);
        piv_.push_back(-1.0);
    }
    const int S = static_cast<int>(p_.size());
    lbeg_.assign(static_cast<size_t>(S) + 1, 0);
    urow_beg_.assign(static_cast<size_t>(S) + 1, 0);
    for (int k = 0; k < m; ++k) {
        const int64_t v = head[static_cast<size_t>(k)];
        if (v >= A.n) continue;
        for (int64_t t = A.ptr[v]; t < A.ptr[v + 1]; ++t) {
            const int st = srow_step[static_cast<size_t>(A.idx[t])];
            if (st >= 0 && A.val[t] != 0.0) ++urow_beg_[static_cast<size_t>(st) + 1];
        }
    }
    for (int st = 0; st < S; ++st) urow_beg_[static_cast<size_t>(st) + 1] += urow_beg_[static_cast<size_t>(st)];
    uidx_.resize(static_cast<size_t>(urow_beg_[static_cast<size_t>(S)]));
    uval_.resize(uidx_.size());
    std::vector<int> ucur(urow_beg_.begin(), urow_beg_.end() - 1);
    // active submatrix: values by column, patterns by row
    std::vector<std::vector<std::pair<int, double>>> col(static_cast<size_t>(m));
    std::vector<std::vector<int>> row(static_cast<size_t>(m));
    std::vector<double> orig_max(static_cast<size_t>(m), 0.0);
    for (int k = 0; k < m; ++k) {
        const int64_t v = head[static_cast<size_t>(k)];
        if (v >= A.n) continue;
        auto &c = col[static_cast<size_t>(k)];
        for (int64_t t = A.ptr[v]; t < A.ptr[v + 1]; ++t) {
            if (A.val[t] == 0.0) continue;
            const int i = static_cast<int>(A.idx[t]);
            orig_max[static_cast<size_t>(k)] = std::max(orig_max[static_cast<size_t>(k)], std::fabs(A.val[t]));
            const int st = srow_step[static_cast<size_t>(i)];
            if (st >= 0) {
                const int e = ucur[static_cast<size_t>(st)]++;
                uidx_[static_cast<size_t>(e)] = k;
                uval_[static_cast<size_t>(e)] = A.val[t];
            } else {
                c.emplace_back(i, A.val[t]);
            }
        }
        for (const auto &e : c) row[static_cast<size_t>(e.first)].push_back(k);
    }
    Buckets cb, rb;
    cb.init(m, m);
    rb.init(m, m);
    std::vector<char> col_live(static_cast<size_t>(m), 0), row_live(static_cast<size_t>(m), 0);
    int live = 0;
    for (int k = 0; k < m; ++k)
        if (head[static_cast<size_t>(k)] < A.n) {
            cb.insert(k, static_cast<int>(col[static_cast<size_t>(k)].size()));
            col_live[static_cast<size_t>(k)] = 1;
            ++live;
        }
    for (int i = 0; i < m; ++i)
        if (srow_step[static_cast<size_t>(i)] < 0) {
            rb.insert(i, static_cast<int>(row[static_cast<size_t>(i)].size()));
            row_live[static_cast<size_t>(i)] = 1;
        }
    std::vector<int> slot(static_cast<size_t>(m), -1);
    // a column whose remaining entries have all shrunk below kLuDep times its
    // original largest entry (orig_max) is dependent on the pivoted ones
    auto col_max = [&](int k) {
        double mx = 0.0;
        for (const auto &e : col[static_cast<size_t>(k)]) mx = std::max(mx, std::fabs(e.second));
        return mx;
    };
    auto find_in_col = [&](int k, int i) -> double {
        for (const auto &e : col[static_cast<size_t>(k)])
            if (e.first == i) return e.second;
        return 0.0;
    };
    std::vector<std::pair<int, double>> lcol, urow;
    while (live > 0) {
        // dependent columns: no entry left
        while (cb.head[0] >= 0) {
            const int k = cb.head[0];
            cb.remove(k);
            col_live[static_cast<size_t>(k)] = 0;
            bad_pos.push_back(k);
            --live;
        }
        if (live == 0) break;
        // Markowitz search with threshold pivoting
        int bp = -1, bq = -1;
        double bcost = std::numeric_limits<double>::infinity(), bmag = 0.0;
        int seen = 0;
        for (int c = 1; c <= m; ++c) {
            for (int k = cb.head[c]; k >= 0; k = cb.next[k]) {
                const double mx = col_max(k);
                const double floor = std::max(kLuTiny, kLuDep * orig_max[static_cast<size_t>(k)]);
                for (const auto &e : col[static_cast<size_t>(k)]) {
                    const double a = std::fabs(e.second);
                    if (a < kLuThreshold * mx || a < floor) continue;
                    const double cost = static_cast<double>(rb.cnt[e.first] - 1) * (c - 1);
                    if (cost < bcost || (cost == bcost && a > bmag)) {
                        bcost = cost; bmag = a; bp = e.first; bq = k;
                    }
                }
                if (bp >= 0 && ++seen >= kSearch) break;
            }
            if (bp >= 0 && seen >= kSearch) break;
            for (int i = rb.head[c]; i >= 0; i = rb.next[i]) {
                for (int k : row[static_cast<size_t>(i)]) {
                    const double a = std::fabs(find_in_col(k, i));
                    if (a < std::max(kLuTiny, kLuDep * orig_max[static_cast<size_t>(k)]) ||
                        a < kLuThreshold * col_max(k))
                        continue;
                    const double cost = static_cast<double>(c - 1) * (cb.cnt[k] - 1);
                    if (cost < bcost || (cost == bcost && a > bmag)) {
                        bcost = cost; bmag = a; bp = i; bq = k;
                    }
                }
                if (bp >= 0 && ++seen >= kSearch) break;
            }
            if (bp >= 0 && (seen >= kSearch || bcost <= static_cast<double>(c) * c)) break;
        }
        if (bp < 0) {
            // every live column is numerically zero: all dependent
            for (int c = 1; c <= m; ++c)
                while (cb.head[c] >= 0) {
                    const int k = cb.head[c];
                    cb.remove(k);
                    col_live[static_cast<size_t>(k)] = 0;
                    bad_pos.push_back(k);
                    --live;
                }
            break;
        }
        // eliminate with pivot (bp, bq)
        const double piv = find_in_col(bq, bp);
        lcol.clear();
        for (const auto &e : col[static_cast<size_t>(bq)]) {
            auto &rr = row[static_cast<size_t>(e.first)];
            rr.erase(std::find(rr.begin(), rr.end(), bq));
            if (e.first != bp) lcol.emplace_back(e.first, e.second / piv);
        }
        col[static_cast<size_t>(bq)].clear();
        cb.remove(bq);
        col_live[static_cast<size_t>(bq)] = 0;
        urow.clear();
        for (int k : row[static_cast<size_t>(bp)]) {
            auto &cc = col[static_cast<size_t>(k)];
            for (size_t t = 0; t < cc.size(); ++t)
                if (cc[t].first == bp) {
                    urow.emplace_back(k, cc[t].second);
                    cc[t] = cc.back();
                    cc.pop_back();
                    break;
                }
        }
        row[static_cast<size_t>(bp)].clear();
        rb.remove(bp);
        row_live[static_cast<size_t>(bp)] = 0;
        // rank-one update of the remaining rows of column bq
        for (const auto &ue : urow) {
            auto &cc = col[static_cast<size_t>(ue.first)];
            for (size_t t = 0; t < cc.size(); ++t) slot[static_cast<size_t>(cc[t].first)] = static_cast<int>(t);
            for (const auto &le : lcol) {
                const double d = le.second * ue.second;
                const int s = slot[static_cast<size_t>(le.first)];
                if (s >= 0) {
                    cc[static_cast<size_t>(s)].second -= d;
                } else {
                    slot[static_cast<size_t>(le.first)] = static_cast<int>(cc.size());
                    cc.emplace_back(le.first, -d);
                    row[static_cast<size_t>(le.first)].push_back(ue.first);
                }
            }
            for (const auto &e : cc) slot[static_cast<size_t>(e.first)] = -1;
        }
        for (const auto &ue : urow) cb.move(ue.first, static_cast<int>(col[static_cast<size_t>(ue.first)].size()));
        for (const auto &le : lcol) rb.move(le.first, static_cast<int>(row[static_cast<size_t>(le.first)].size()));
        p_.push_back(bp);
        q_.push_back(bq);
        piv_.push_back(piv);
        for (const auto &le : lcol) { lidx_.push_back(le.first); lval_.push_back(le.second); }
        lbeg_.push_back(static_cast<int>(lidx_.size()));
        for (const auto &ue : urow) { uidx_.push_back(ue.first); uval_.push_back(ue.second); }
        urow_beg_.push_back(static_cast<int>(uidx_.size()));
        --live;
    }
    for (int i = 0; i < m; ++i)
        if (row_live[static_cast<size_t>(i)]) free_rows.push_back(i);
    // U by columns: the entry (position q', u) of step j's row lies in the
    // column of the step that pivoted q'.  Only a complete factorisation is
    // ever solved with; an incomplete one (dependent positions, which have
    // no step) is repaired and factorised again, so it gets no column form.
    if (!bad_pos.empty()) {
        ucol_beg_.clear();
        ucol_row_.clear();
        ucol_val_.clear();
        return;
    }
    const int K = static_cast<int>(p_.size());
    std::vector<int> step_of_pos(static_cast<size_t>(m), -1);
    for (int k = 0; k < K; ++k) step_of_pos[static_cast<size_t>(q_[k])] = k;
    ucol_beg_.assign(static_cast<size_t>(K) + 1, 0);
    for (size_t e = 0; e < uidx_.size(); ++e) ++ucol_beg_[static_cast<size_t>(step_of_pos[static_cast<size_t>(uidx_[e])]) + 1];
    for (int k = 0; k < K; ++k) ucol_beg_[static_cast<size_t>(k) + 1] += ucol_beg_[static_cast<size_t>(k)];
    ucol_row_.resize(uidx_.size());
    ucol_val_.resize(uidx_.size());
    std::vector<int> cur(ucol_beg_.begin(), ucol_beg_.end() - 1);
    for (int j = 0; j < K; ++j)
        for (int e = urow_beg_[j]; e < urow_beg_[j + 1]; ++e) {
            const int c = cur[static_cast<size_t>(step_of_pos[static_cast<size_t>(uidx_[e])])]++;
            ucol_row_[static_cast<size_t>(c)] = p_[j];
            ucol_val_[static_cast<size_t>(c)] = uval_[e];
        }
}

void Factor::ftran(std::vector<double> &v) const {
    const int K = static_cast<int>(p_.size());
    for (int k = 0; k < K; ++k) {
        const double t = v[static_cast<size_t>(p_[k])];
        if (t == 0.0) continue;
        for (int e = lbeg_[k]; e < lbeg_[k + 1]; ++e) v[static_cast<size_t>(lidx_[e])] -= lval_[e] * t;
    }
    // U x = w by columns, last step first: each unknown, once known, is
    // removed from the rows of the earlier steps; a zero unknown costs one test
    w_.resize(static_cast<size_t>(m_));
    for (int k = K - 1; k >= 0; --k) {
        const double xk = v[static_cast<size_t>(p_[k])] / piv_[k];
        w_[static_cast<size_t>(q_[k])] = xk;
        if (xk == 0.0) continue;
        for (int e = ucol_beg_[k]; e < ucol_beg_[k + 1]; ++e) v[static_cast<size_t>(ucol_row_[e])] -= ucol_val_[e] * xk;
    }
    v.swap(w_);
    for (size_t t = 0; t < eta_r_.size(); ++t) {
        const int r = eta_r_[t];
        const double xr = v[static_cast<size_t>(r)] / eta_piv_[t];
        v[static_cast<size_t>(r)] = xr;
        if (xr == 0.0) continue;
        for (int e = eta_beg_[t]; e < eta_beg_[t + 1]; ++e) v[static_cast<size_t>(eta_idx_[e])] -= eta_val_[e] * xr;
    }
}

void Factor::btran(std::vector<double> &v) const {
    for (size_t t = eta_r_.size(); t-- > 0;) {
        const int r = eta_r_[t];
        double s = v[static_cast<size_t>(r)];
        for (int e = eta_beg_[t]; e < eta_beg_[t + 1]; ++e) s -= eta_val_[e] * v[static_cast<size_t>(eta_idx_[e])];
        v[static_cast<size_t>(r)] = s / eta_piv_[t];
    }
    const int K = static_cast<int>(p_.size());
    w_.assign(static_cast<size_t>(m_), 0.0);
    for (int k = 0; k < K; ++k) {
        const double z = v[static_cast<size_t>(q_[k])] / piv_[k];
        w_[static_cast<size_t>(p_[k])] = z;
        if (z == 0.0) continue;
        for (int e = urow_beg_[k]; e < urow_beg_[k + 1]; ++e) v[static_cast<size_t>(uidx_[e])] -= uval_[e] * z;
    }
    for (int k = K - 1; k >= 0; --k) {
        double s = w_[static_cast<size_t>(p_[k])];
        for (int e = lbeg_[k]; e < lbeg_[k + 1]; ++e) s -= lval_[e] * w_[static_cast<size_t>(lidx_[e])];
        w_[static_cast<size_t>(p_[k])] = s;
    }
    v.swap(w_);
}

void Factor::update(int r, const std::vector<double> &alpha, const std::vector<int> &nz) {
    eta_r_.push_back(r);
    eta_piv_.push_back(alpha[static_cast<size_t>(r)]);
    for (int i : nz)
        if (i != r) {
            eta_idx_.push_back(i);
            eta_val_.push_back(alpha[static_cast<size_t>(i)]);
        }
    eta_beg_.push_back(static_cast<int>(eta_idx_.size()));
}

// ---- the simplex ------------------------------------------------------------
class Simplex {
  public:
    Simplex(const mpdhg_lp &lp, const mpdhg_polish_opts &o)
        : A_{lp.m, lp.n, lp.col_ptr, lp.row_ind, lp.vals}, lp_(lp), o_(o),
          m_(static_cast<int>(lp.m)), n_(lp.n), N_(lp.n + lp.m), f_(static_cast<int>(lp.m)) {
        lb_.resize(static_cast<size_t>(N_));
        ub_.resize(static_cast<size_t>(N_));
        c_.assign(static_cast<size_t>(N_), 0.0);
        for (int64_t j = 0; j < n_; ++j) {
            lb_[static_cast<size_t>(j)] = finite_lo(lp.col_lower[j]);
            ub_[static_cast<size_t>(j)] = finite_hi(lp.col_upper[j]);
            c_[static_cast<size_t>(j)] = lp.c[j];
        }
        for (int64_t i = 0; i < lp.m; ++i) {
            lb_[static_cast<size_t>(n_ + i)] = finite_lo(lp.row_lower[i]);
            ub_[static_cast<size_t>(n_ + i)] = finite_hi(lp.row_upper[i]);
        }
        // rows of A for the pricing products a_j^T y over all columns
        rptr_.assign(static_cast<size_t>(m_ + 1), 0);
        for (int64_t t = 0; t < lp.col_ptr[n_]; ++t) ++rptr_[static_cast<size_t>(lp.row_ind[t] + 1)];
        for (int i = 0; i < m_; ++i) rptr_[static_cast<size_t>(i + 1)] += rptr_[static_cast<size_t>(i)];
        rcol_.resize(static_cast<size_t>(lp.col_ptr[n_]));
        rval_.resize(static_cast<size_t>(lp.col_ptr[n_]));
        std::vector<int64_t> fill(rptr_.begin(), rptr_.end() - 1);
        for (int64_t j = 0; j < n_; ++j)
            for (int64_t t = lp.col_ptr[j]; t < lp.col_ptr[j + 1]; ++t) {
                const int64_t s = fill[static_cast<size_t>(lp.row_ind[t])]++;
                rcol_[static_cast<size_t>(s)] = j;
                rval_[static_cast<size_t>(s)] = lp.vals[t];
            }
    }

    // Warm start (primal_simplex.cpp:77-148): basic variables are placed in
    // column-then-row order; nonbasic ones sit at the bound their status
    // names when it is finite, else at a finite bound, else at zero (free).
    // Returns false (cold start) when the basis does not hold m variables.
    //
    // With a start point x0 (columns; the row variables take A x0), a
    // nonbasic variable that x0 puts strictly inside its bounds keeps that
    // value as a superbasic one until the push phase moves it.
    bool start(const int8_t *col_basis, const int8_t *row_basis, const double *x0) {
        if (x0) {
            xs_.assign(static_cast<size_t>(N_), 0.0);
            for (int64_t j = 0; j < n_; ++j) {
                xs_[static_cast<size_t>(j)] = x0[j];
                for (int64_t t = lp_.col_ptr[j]; t < lp_.col_ptr[j + 1]; ++t)
                    xs_[static_cast<size_t>(n_ + lp_.row_ind[t])] += lp_.vals[t] * x0[j];
            }
        }
        head_.clear();
        st_.assign(static_cast<size_t>(N_), kLower);
        pos_.assign(static_cast<size_t>(N_), -1);
        x_.assign(static_cast<size_t>(N_), 0.0);
        if (x0) {
            // Crossover start: the slack basis (well conditioned, and the row
            // activities A x0 are its basic values), every column at x0 -- at
            // a bound, or superbasic inside -- and the push order: the columns
            // the given basis names basic first, each group most interior first.
            for (int i = 0; i < m_; ++i) {
                head_.push_back(n_ + i);
                pos_[static_cast<size_t>(n_ + i)] = i;
                st_[static_cast<size_t>(n_ + i)] = kBasic;
            }
            prio_.assign(static_cast<size_t>(N_), 0);
            for (int64_t j = 0; j < n_; ++j) {
                const int8_t b = col_basis ? col_basis[j] : static_cast<int8_t>(MPDHG_BASIS_AT_LOWER);
                place_nonbasic(j, b == MPDHG_BASIS_AT_UPPER ? kUpper : kLower);
                keep_inside(j);
                prio_[static_cast<size_t>(j)] = (col_basis && b == MPDHG_BASIS_BASIC) ? 1 : 0;
            }
            warm_ = col_basis && row_basis;
            refactor();
            return warm_;
        }
        bool warm = col_basis && row_basis;
        if (warm) {
            for (int64_t v = 0; v < N_; ++v)
                if ((v < n_ ? col_basis[v] : row_basis[v - n_]) == MPDHG_BASIS_BASIC) head_.push_back(v);
            if (static_cast<int64_t>(head_.size()) != m_) {
                warm = false;
                head_.clear();
            }
        }
        if (!warm)
            for (int i = 0; i < m_; ++i) head_.push_back(n_ + i);
        for (int k = 0; k < m_; ++k) {
            pos_[static_cast<size_t>(head_[static_cast<size_t>(k)])] = k;
            st_[static_cast<size_t>(head_[static_cast<size_t>(k)])] = kBasic;
        }
        for (int64_t v = 0; v < N_; ++v) {
            if (pos_[static_cast<size_t>(v)] >= 0) continue;
            const int8_t want = warm ? (v < n_ ? col_basis[v] : row_basis[v - n_]) : static_cast<int8_t>(kLower);
            place_nonbasic(v, want == MPDHG_BASIS_AT_UPPER ? kUpper : kLower);
            if (warm) keep_inside(v);
        }
        warm_ = warm;
        refactor();
        return warm;
    }

    // The pivoting loop; a factorisation that fails mid-solve ends it with
    // MPDHG_STATUS_NUMERICAL_FAILURE.
    int run(double t0);
    void results(mpdhg_polish_result &r) const;

  private:
    int iterate(double t0);
    static double finite_lo(double v) { return (std::isinf(v) || v <= -kBig) ? -HUGE_VAL : v; }
    static double finite_hi(double v) { return (std::isinf(v) || v >= kBig) ? HUGE_VAL : v; }
    bool has_lo(int64_t v) const { return std::isfinite(lb_[static_cast<size_t>(v)]); }
    bool has_hi(int64_t v) const { return std::isfinite(ub_[static_cast<size_t>(v)]); }

    // A nonbasic variable whose start value lies inside its bounds by more
    // than the primal tolerance stays there (superbasic).
    void keep_inside(int64_t v) {
        if (xs_.empty()) return;
        const size_t u = static_cast<size_t>(v);
        const double xv = xs_[u], tp = o_.primal_tol;
        if (!std::isfinite(xv) || !(xv > lb_[u] + tp) || !(xv < ub_[u] - tp)) return;
        if (!has_lo(v) && !has_hi(v) && std::fabs(xv) <= tp) return;
        st_[u] = kSuper;
        x_[u] = xv;
    }

    // Primal push: each superbasic variable moves toward a bound (the one
    // its reduced cost prefers, else the nearer; a free one toward zero)
    // until it gets there and becomes nonbasic, or a basic variable reaches
    // one of its bounds first, leaves, and the pushed variable takes its
    // place.  No pricing: one FTRAN and one ratio test per step.
    void push() {
        const double tp = o_.primal_tol, td = o_.dual_tol;
        std::vector<int64_t> todo;
        std::vector<double> depth(static_cast<size_t>(N_), 0.0);
        for (int64_t v = 0; v < N_; ++v) {
            const size_t u = static_cast<size_t>(v);
            if (st_[u] != kSuper) continue;
            todo.push_back(v);
            depth[u] = std::min(has_lo(v) ? x_[u] - lb_[u] : HUGE_VAL, has_hi(v) ? ub_[u] - x_[u] : HUGE_VAL);
        }
        if (todo.empty()) return;
        std::stable_sort(todo.begin(), todo.end(), [&](int64_t a, int64_t b) {
            const int pa = prio_.empty() ? 0 : prio_[static_cast<size_t>(a)];
            const int pb = prio_.empty() ? 0 : prio_[static_cast<size_t>(b)];
            if (pa != pb) return pa > pb;
            return depth[static_cast<size_t>(a)] > depth[static_cast<size_t>(b)];
        });
        std::vector<double> cb(static_cast<size_t>(m_)), alpha;
        for (int k = 0; k < m_; ++k) cb[static_cast<size_t>(k)] = c_[static_cast<size_t>(head_[static_cast<size_t>(k)])];
        reduced(cb, true);
        for (const int64_t v : todo) {
            const size_t u = static_cast<size_t>(v);
            if (st_[u] != kSuper) continue;
            const double xv = x_[u];
            int dir;
            double own;
            if (!has_lo(v) && !has_hi(v)) {
                dir = xv > 0 ? -1 : 1;
                own = std::fabs(xv);
            } else {
                if (d_[u] < -td) dir = 1;
                else if (d_[u] > td) dir = -1;
                else dir = (has_hi(v) && (!has_lo(v) || ub_[u] - xv < xv - lb_[u])) ? 1 : -1;
                if (dir > 0 && !has_hi(v)) dir = -1;
                if (dir < 0 && !has_lo(v)) dir = 1;
                own = dir > 0 ? ub_[u] - xv : xv - lb_[u];
            }
            column(v, alpha);
            f_.ftran(alpha);
            nonzeros(alpha);
            double tmax = own;
            for (const int k : nz_) {
                const double a = alpha[static_cast<size_t>(k)];
                if (std::fabs(a) < kPivTol) continue;
                const double rate = -dir * a;
                const size_t b = static_cast<size_t>(head_[static_cast<size_t>(k)]);
                // a basic variable outside a bound counts as on it: the push
                // never moves it further out
                if (rate < 0 && std::isfinite(lb_[b]))
                    tmax = std::min(tmax, (std::max(0.0, x_[b] - lb_[b]) + tp) / -rate);
                else if (rate > 0 && std::isfinite(ub_[b]))
                    tmax = std::min(tmax, (std::max(0.0, ub_[b] - x_[b]) + tp) / rate);
            }
            int r = -1;
            double rmag = 0.0, tstep = own;
            bool to_lower = false;
            for (const int k : nz_) {
                const double a = alpha[static_cast<size_t>(k)];
                if (std::fabs(a) < kPivTol) continue;
                const double rate = -dir * a;
                const size_t b = static_cast<size_t>(head_[static_cast<size_t>(k)]);
                double t = HUGE_VAL;
                bool low = false;
                if (rate < 0 && std::isfinite(lb_[b])) { t = std::max(0.0, x_[b] - lb_[b]) / -rate; low = true; }
                else if (rate > 0 && std::isfinite(ub_[b])) t = std::max(0.0, ub_[b] - x_[b]) / rate;
                if (t <= tmax && std::fabs(a) > rmag) {
                    rmag = std::fabs(a);
                    r = k;
                    tstep = t;
                    to_lower = low;
                }
            }
            ++pushes_;
            if (r < 0 || own <= tstep) {
                for (const int k : nz_)
                    x_[static_cast<size_t>(head_[static_cast<size_t>(k)])] -= dir * own * alpha[static_cast<size_t>(k)];
                if (!has_lo(v) && !has_hi(v)) {
                    x_[u] = 0.0;
                    st_[u] = kFree;
                } else {
                    x_[u] = dir > 0 ? ub_[u] : lb_[u];
                    st_[u] = dir > 0 ? kUpper : kLower;
                }
                continue;
            }
            const double t = std::max(0.0, tstep);
            for (const int k : nz_)
                x_[static_cast<size_t>(head_[static_cast<size_t>(k)])] -= dir * t * alpha[static_cast<size_t>(k)];
            x_[u] += dir * t;
            const int64_t leave = head_[static_cast<size_t>(r)];
            const size_t ul = static_cast<size_t>(leave);
            pos_[ul] = -1;
            st_[ul] = to_lower ? kLower : kUpper;
            x_[ul] = to_lower ? lb_[ul] : ub_[ul];
            head_[static_cast<size_t>(r)] = v;
            pos_[u] = r;
            st_[u] = kBasic;
            f_.update(r, alpha, nz_);
            if (f_.updates() >= o_.refactor_every) refactor();
        }
        refactor();
    }

    void place_nonbasic(int64_t v, Stat want) {
        const size_t u = static_cast<size_t>(v);
        if (want == kUpper && has_hi(v)) {
            st_[u] = kUpper;
            x_[u] = ub_[u];
        } else if (has_lo(v)) {
            st_[u] = kLower;
            x_[u] = lb_[u];
        } else if (has_hi(v)) {
            st_[u] = kUpper;
            x_[u] = ub_[u];
        } else {
            st_[u] = kFree;
            x_[u] = 0.0;
        }
    }

    // Factorise the basis, repairing dependent columns with slacks; then
    // recompute the basic values from the nonbasic ones.
    void refactor() {
        d_valid_ = false;
        std::vector<int> bad, freer;
        for (int pass = 0; pass < 3; ++pass) {
            f_.factorize(A_, head_, bad, freer);
            ++refactors_;
            if (bad.empty()) break;
            for (size_t t = 0; t < bad.size(); ++t) {
                const int k = bad[t];
                const int64_t old = head_[static_cast<size_t>(k)];
                const int64_t slack = n_ + freer[t];
                pos_[static_cast<size_t>(old)] = -1;
                place_nonbasic(old, (has_hi(old) && std::isfinite(x_[static_cast<size_t>(old)]) &&
                                     x_[static_cast<size_t>(old)] >= ub_[static_cast<size_t>(old)])
                                        ? kUpper : kLower);
                keep_inside(old);
                head_[static_cast<size_t>(k)] = slack;
                pos_[static_cast<size_t>(slack)] = k;
                st_[static_cast<size_t>(slack)] = kBasic;
                ++repairs_;
            }
        }
        if (!bad.empty()) throw FactorError();
        basic_values();
    }

    // x_B = B^-1 (-N x_N): the row variables contribute +x (column -e_i).
    void basic_values() {
        std::vector<double> rhs(static_cast<size_t>(m_), 0.0);
        for (int64_t j = 0; j < n_; ++j) {
            if (pos_[static_cast<size_t>(j)] >= 0) continue;
            const double xj = x_[static_cast<size_t>(j)];
            if (xj == 0.0) continue;
            for (int64_t t = lp_.col_ptr[j]; t < lp_.col_ptr[j + 1]; ++t)
                rhs[static_cast<size_t>(lp_.row_ind[t])] -= lp_.vals[t] * xj;
        }
        for (int i = 0; i < m_; ++i)
            if (pos_[static_cast<size_t>(n_ + i)] < 0) rhs[static_cast<size_t>(i)] += x_[static_cast<size_t>(n_ + i)];
        f_.ftran(rhs);
        for (int k = 0; k < m_; ++k) x_[static_cast<size_t>(head_[static_cast<size_t>(k)])] = rhs[static_cast<size_t>(k)];
    }

    // The positions where an FTRAN result is nonzero, ascending: the ratio
    // tests, the updates of the basic values and the eta file visit only these.
    void nonzeros(const std::vector<double> &a) {
        nz_.clear();
        for (int k = 0; k < m_; ++k)
            if (a[static_cast<size_t>(k)] != 0.0) nz_.push_back(k);
    }

    void column(int64_t v, std::vector<double> &out) const {
        out.assign(static_cast<size_t>(m_), 0.0);
        if (v >= n_) {
            out[static_cast<size_t>(v - n_)] = -1.0;
        } else {
            for (int64_t t = lp_.col_ptr[v]; t < lp_.col_ptr[v + 1]; ++t)
                out[static_cast<size_t>(lp_.row_ind[t])] = lp_.vals[t];
        }
    }

    // Reduced costs of every variable for basic costs cb (by position) and
    // nonbasic costs cn (the phase's own: c in phase 2, 0 in phase 1).
    void reduced(const std::vector<double> &cb, bool phase2) {
        y_ = cb;
        f_.btran(y_);
        d_.assign(static_cast<size_t>(N_), 0.0);
        for (int64_t j = 0; j < n_; ++j) d_[static_cast<size_t>(j)] = phase2 ? c_[static_cast<size_t>(j)] : 0.0;
        for (int i = 0; i < m_; ++i) {
            const double yi = y_[static_cast<size_t>(i)];
            d_[static_cast<size_t>(n_ + i)] = yi;
            if (yi == 0.0) continue;
            for (int64_t t = rptr_[static_cast<size_t>(i)]; t < rptr_[static_cast<size_t>(i + 1)]; ++t)
                d_[static_cast<size_t>(rcol_[static_cast<size_t>(t)])] -= rval_[static_cast<size_t>(t)] * yi;
        }
        for (int k = 0; k < m_; ++k) d_[static_cast<size_t>(head_[static_cast<size_t>(k)])] = 0.0;
    }

    const Cols A_;
    const mpdhg_lp &lp_;
    const mpdhg_polish_opts &o_;
    const int m_;
    const int64_t n_, N_;
    Factor f_;
    std::vector<double> lb_, ub_, c_, x_, y_, d_;
    std::vector<int64_t> head_, pos_;
    std::vector<int8_t> st_;
    std::vector<int64_t> rptr_, rcol_;
    std::vector<double> rval_;
    // Devex: after a pivot on row r with entering q, every nonbasic weight
    // becomes max(w_j, (alpha_rj / alpha_rq)^2 w_q), and the leaving
    // variable gets max(w_q / alpha_rq^2, 1); the pivot row comes from
    // rho = B^-T e_r (before the basis change).
    void devex(int r, int64_t q, double arq) {
        rho_.assign(static_cast<size_t>(m_), 0.0);
        rho_[static_cast<size_t>(r)] = 1.0;
        f_.btran(rho_);
        const double wq = w_[static_cast<size_t>(q)];
        row_.assign(static_cast<size_t>(n_), 0.0);
        for (int i = 0; i < m_; ++i) {
            const double ri = rho_[static_cast<size_t>(i)];
            if (ri == 0.0) continue;
            for (int64_t t = rptr_[static_cast<size_t>(i)]; t < rptr_[static_cast<size_t>(i + 1)]; ++t)
                row_[static_cast<size_t>(rcol_[static_cast<size_t>(t)])] += rval_[static_cast<size_t>(t)] * ri;
        }
        double wmax = 0.0;
        for (int64_t v = 0; v < N_; ++v) {
            const size_t u = static_cast<size_t>(v);
            if (st_[u] == kBasic || v == q) continue;
            const double a = v < n_ ? row_[u] : -rho_[static_cast<size_t>(v - n_)];
            if (a == 0.0) continue;
            const double ratio = a / arq;
            w_[u] = std::max(w_[u], ratio * ratio * wq);
            wmax = std::max(wmax, w_[u]);
        }
        w_[static_cast<size_t>(head_[static_cast<size_t>(r)])] = std::max(wq / (arq * arq), 1.0);
        if (wmax > 1e6) w_.assign(static_cast<size_t>(N_), 1.0);  // a new reference framework
    }

    // Expand every non-fixed finite bound by a small, variable-dependent
    // amount so that ratio-test ties (degenerate vertices) split; the true
    // bounds come back at the end and the loop confirms optimality there.
    void perturb() {
        lb0_ = lb_;
        ub0_ = ub_;
        for (int64_t v = 0; v < N_; ++v) {
            const size_t u = static_cast<size_t>(v);
            if (!(lb_[u] < ub_[u])) continue;
            const uint64_t h = (static_cast<uint64_t>(v) + 1) * 0x9E3779B97F4A7C15ull;
            const double r1 = static_cast<double>((h >> 11) & 0xFFFFF) / 1048576.0;
            const double r2 = static_cast<double>((h >> 33) & 0xFFFFF) / 1048576.0;
            if (std::isfinite(lb_[u])) lb_[u] -= kPerturb * (1.0 + r1) * std::max(1.0, std::fabs(lb_[u]));
            if (std::isfinite(ub_[u])) ub_[u] += kPerturb * (1.0 + r2) * std::max(1.0, std::fabs(ub_[u]));
            if (st_[u] == kLower && std::isfinite(lb_[u])) x_[u] = lb_[u];
            if (st_[u] == kUpper && std::isfinite(ub_[u])) x_[u] = ub_[u];
        }
        perturbed_ = true;
        basic_values();
    }
    void unperturb() {
        lb_ = lb0_;
        ub_ = ub0_;
        for (int64_t v = 0; v < N_; ++v) {
            const size_t u = static_cast<size_t>(v);
            if (st_[u] == kLower && std::isfinite(lb_[u])) x_[u] = lb_[u];
            if (st_[u] == kUpper && std::isfinite(ub_[u])) x_[u] = ub_[u];
        }
        perturbed_ = false;
        refactor();
    }

    std::vector<double> w_, rho_, row_, lb0_, ub0_, xs_;
    std::vector<int8_t> prio_;
    std::vector<int> nz_;
    int64_t pushes_ = 0;
    bool perturbed_ = false;
    int64_t degenerate_ = 0;
    bool d_valid_ = false;  // d_ and y_ match the current basis (phase 2)
    bool warm_ = false;
    int64_t iters_ = 0, refactors_ = 0, repairs_ = 0, flips_ = 0;
    int status_ = MPDHG_STATUS_UNKNOWN;
    double infeas_ = 0.0;
};

int Simplex::run(double t0) {
    try {
        return iterate(t0);
    } catch (const FactorError &) {
        return status_ = MPDHG_STATUS_NUMERICAL_FAILURE;
    }
}

int Simplex::iterate(double t0) {
    const double tp = o_.primal_tol, td = o_.dual_tol;
    const int64_t cap = o_.max_iterations > 0 ? o_.max_iterations : 10 * N_ + 10000;
    std::vector<double> cb(static_cast<size_t>(m_)), alpha;
    w_.assign(static_cast<size_t>(N_), 1.0);
    push();
    perturb();
    for (;;) {
        if (iters_ >= cap) return status_ = MPDHG_STATUS_ITERATION_LIMIT;
        if (o_.time_limit_sec > 0 && now() - t0 > o_.time_limit_sec) return status_ = MPDHG_STATUS_TIME_LIMIT;
        // phase 1 while a basic variable violates a bound by more than tp:
        // cost -1 below the lower bound, +1 above the upper one, 0 otherwise
        bool phase2 = true;
        infeas_ = 0.0;
        for (int k = 0; k < m_; ++k) {
            const size_t v = static_cast<size_t>(head_[static_cast<size_t>(k)]);
            const double below = lb_[v] - x_[v], above = x_[v] - ub_[v];
            if (below > tp) infeas_ += below;
            if (above > tp) infeas_ += above;
            if (below > tp || above > tp) phase2 = false;
        }
        for (int k = 0; k < m_; ++k) {
            const size_t v = static_cast<size_t>(head_[static_cast<size_t>(k)]);
            cb[static_cast<size_t>(k)] = phase2 ? c_[v] : (x_[v] < lb_[v] - tp ? -1.0 : (x_[v] > ub_[v] + tp ? 1.0 : 0.0));
        }
        // phase 2 keeps the reduced costs up to date across pivots (below);
        // phase 1, and the first phase-2 iteration after a factorisation,
        // compute them afresh
        if (!(phase2 && d_valid_)) reduced(cb, phase2);
        d_valid_ = phase2;
        // pricing: the largest d_j^2 / w_j over the dual infeasible variables
        // (Devex reference weights w); a fixed variable never enters
        int64_t q = -1;
        double best = 0.0;
        int dir = 0;
        for (int64_t v = 0; v < N_; ++v) {
            const size_t u = static_cast<size_t>(v);
            if (st_[u] == kBasic || lb_[u] == ub_[u]) continue;
            const double dv = d_[u];
            int vdir = 0;
            if (st_[u] == kLower) {
                if (dv < -td) vdir = 1;
            } else if (st_[u] == kUpper) {
                if (dv > td) vdir = -1;
            } else if (std::fabs(dv) > td) {
                vdir = dv < 0 ? 1 : -1;
            }
            if (!vdir) continue;
            const double score = dv * dv / w_[u];
            if (score > best) {
                best = score;
                q = v;
                dir = vdir;
            }
        }
        if (q < 0) {
            if (f_.updates() > 0) {  // confirm on a fresh factorisation
                refactor();
                continue;
            }
            if (phase2 && perturbed_) {  // back to the true bounds, then confirm
                unperturb();
                continue;
            }
            return status_ = phase2 ? MPDHG_STATUS_OPTIMAL : MPDHG_STATUS_INFEASIBLE;
        }
        ++iters_;
        column(q, alpha);
        f_.ftran(alpha);
        nonzeros(alpha);
        // ratio test: pass 1 with the tolerance, pass 2 the largest |alpha|
        const size_t uq = static_cast<size_t>(q);
        const double flip = (has_lo(q) && has_hi(q)) ? ub_[uq] - lb_[uq] : HUGE_VAL;
        double tmax = flip;
        for (const int k : nz_) {
            const double a = alpha[static_cast<size_t>(k)];
            if (std::fabs(a) < kPivTol) continue;
            const double rate = -dir * a;  // d x_B[k] / d t
            const size_t v = static_cast<size_t>(head_[static_cast<size_t>(k)]);
            const double xv = x_[v];
            double lim = HUGE_VAL;
            if (rate < 0) {
                if (xv > ub_[v] + tp) lim = (xv - ub_[v]) / -rate;          // infeasible above: first breakpoint
                else if (xv >= lb_[v] - tp && has_lo(static_cast<int64_t>(v))) lim = (xv - lb_[v] + tp) / -rate;
            } else {
                if (xv < lb_[v] - tp) lim = (lb_[v] - xv) / rate;
                else if (xv <= ub_[v] + tp && has_hi(static_cast<int64_t>(v))) lim = (ub_[v] - xv + tp) / rate;
            }
            tmax = std::min(tmax, lim);
        }
        if (!std::isfinite(tmax)) return status_ = phase2 ? MPDHG_STATUS_UNBOUNDED : MPDHG_STATUS_NUMERICAL_FAILURE;
        int r = -1;
        double rmag = 0.0, tstep = 0.0;
        bool to_lower = false;
        for (const int k : nz_) {
            const double a = alpha[static_cast<size_t>(k)];
            if (std::fabs(a) < kPivTol) continue;
            const double rate = -dir * a;
            const size_t v = static_cast<size_t>(head_[static_cast<size_t>(k)]);
            const double xv = x_[v];
            double t = HUGE_VAL;
            bool low = false;
            if (rate < 0) {
                if (xv > ub_[v] + tp) { t = (xv - ub_[v]) / -rate; low = false; }
                else if (xv >= lb_[v] - tp && has_lo(static_cast<int64_t>(v))) { t = (xv - lb_[v]) / -rate; low = true; }
            } else {
                if (xv < lb_[v] - tp) { t = (lb_[v] - xv) / rate; low = true; }
                else if (xv <= ub_[v] + tp && has_hi(static_cast<int64_t>(v))) { t = (ub_[v] - xv) / rate; low = false; }
            }
            if (t <= tmax && std::fabs(a) > rmag) {
                rmag = std::fabs(a);
                r = k;
                tstep = t;
                to_lower = low;
            }
        }
        if (r < 0 || flip <= tstep) {
            // the entering variable reaches its other bound first
            if (!std::isfinite(flip)) return status_ = MPDHG_STATUS_NUMERICAL_FAILURE;
            const double t = flip;
            for (const int k : nz_)
                x_[static_cast<size_t>(head_[static_cast<size_t>(k)])] -= dir * t * alpha[static_cast<size_t>(k)];
            x_[uq] = dir > 0 ? ub_[uq] : lb_[uq];
            st_[uq] = dir > 0 ? kUpper : kLower;
            ++flips_;
            continue;
        }
        const double t = std::max(0.0, tstep);
        if (t <= 1e-12) ++degenerate_;
        devex(r, q, alpha[static_cast<size_t>(r)]);
        if (d_valid_) {
            // y += th rho and d_j -= th alpha_rj (alpha_rj = a_j^T rho, -rho_i for the
            // row variable i), th = d_q / alpha_rq: d_q becomes 0 and the leaving
            // variable's -th
            const double th = d_[uq] / alpha[static_cast<size_t>(r)];
            for (int i = 0; i < m_; ++i) y_[static_cast<size_t>(i)] += th * rho_[static_cast<size_t>(i)];
            for (int64_t v = 0; v < N_; ++v) {
                const size_t u = static_cast<size_t>(v);
                if (st_[u] == kBasic || v == q) continue;
                const double a = v < n_ ? row_[u] : -rho_[static_cast<size_t>(v - n_)];
                if (a != 0.0) d_[u] -= th * a;
            }
            d_[uq] = 0.0;
            d_[static_cast<size_t>(head_[static_cast<size_t>(r)])] = -th;
        }
        for (const int k : nz_)
            x_[static_cast<size_t>(head_[static_cast<size_t>(k)])] -= dir * t * alpha[static_cast<size_t>(k)];
        x_[uq] += dir * t;
        const int64_t leave = head_[static_cast<size_t>(r)];
        const size_t ul = static_cast<size_t>(leave);
        pos_[ul] = -1;
        st_[ul] = to_lower ? kLower : kUpper;
        x_[ul] = to_lower ? lb_[ul] : ub_[ul];
        head_[static_cast<size_t>(r)] = q;
        pos_[uq] = r;
        st_[uq] = kBasic;
        f_.update(r, alpha, nz_);
        if (f_.updates() >= o_.refactor_every) refactor();
    }
}

void Simplex::results(mpdhg_polish_result &r) const {
    r.status = status_;
    r.iterations = iters_;
    r.refactorizations = refactors_;
    r.repaired = repairs_;
    r.bound_flips = flips_;
    r.degenerate_pivots = degenerate_;
    r.push_steps = pushes_;
    r.warm_start = warm_ ? 1 : 0;
    r.primal_infeasibility = infeas_;
    double obj = o_.obj_offset;
    for (int64_t j = 0; j < n_; ++j) obj += lp_.c[j] * x_[static_cast<size_t>(j)];
    r.objective = obj;
    if (r.x)
        for (int64_t j = 0; j < n_; ++j) r.x[j] = x_[static_cast<size_t>(j)];
    if (r.row_activity) {
        for (int i = 0; i < m_; ++i) r.row_activity[i] = 0.0;
        for (int64_t j = 0; j < n_; ++j)
            for (int64_t t = lp_.col_ptr[j]; t < lp_.col_ptr[j + 1]; ++t)
                r.row_activity[lp_.row_ind[t]] += lp_.vals[t] * x_[static_cast<size_t>(j)];
    }
    if (r.y)
        for (int i = 0; i < m_; ++i) r.y[i] = y_[static_cast<size_t>(i)];
    if (r.reduced_costs)
        for (int64_t j = 0; j < n_; ++j) r.reduced_costs[j] = d_[static_cast<size_t>(j)];
    if (r.col_basis_out)
        for (int64_t j = 0; j < n_; ++j) r.col_basis_out[j] = st_[static_cast<size_t>(j)];
    if (r.row_basis_out)
        for (int i = 0; i < m_; ++i) r.row_basis_out[i] = st_[static_cast<size_t>(n_ + i)];
}

}  // namespace

}  // namespace mpdhg

extern "C" {

void mpdhg_polish_opts_default(mpdhg_polish_opts *o) {
    if (!o) return;
    o->max_iterations = -1;
    o->primal_tol = 1e-6;
    o->dual_tol = 1e-6;
    o->obj_offset = 0.0;
    o->refactor_every = 100;
    o->time_limit_sec = 0.0;
}

int mpdhg_polish(const mpdhg_lp *lp, const int8_t *col_basis, const int8_t *row_basis,
                 const double *x_start, const mpdhg_polish_opts *opts, mpdhg_polish_result *res) {
    if (!lp || !res) return mpdhg::fail(MPDHG_E_ARG, "polish: NULL argument");
    if (lp->m < 0 || lp->n < 0 || lp->nnz < 0 || lp->m > (int64_t{1} << 30) || lp->n > (int64_t{1} << 30) ||
        !lp->col_ptr ||
        (lp->nnz > 0 && (!lp->row_ind || !lp->vals)) || (lp->n > 0 && (!lp->c || !lp->col_lower || !lp->col_upper)) ||
        (lp->m > 0 && (!lp->row_lower || !lp->row_upper)))
        return mpdhg::fail(MPDHG_E_ARG, "polish: malformed LP");
    if (lp->col_ptr[0] != 0 || lp->col_ptr[lp->n] != lp->nnz)
        return mpdhg::fail(MPDHG_E_ARG, "polish: col_ptr must start at 0 and end at nnz");
    for (int64_t j = 0; j < lp->n; ++j)
        if (lp->col_ptr[j + 1] < lp->col_ptr[j]) return mpdhg::fail(MPDHG_E_ARG, "polish: col_ptr not monotone");
    for (int64_t t = 0; t < lp->col_ptr[lp->n]; ++t)
        if (lp->row_ind[t] < 0 || lp->row_ind[t] >= lp->m) return mpdhg::fail(MPDHG_E_ARG, "polish: row index out of range");
    mpdhg_polish_opts o;
    mpdhg_polish_opts_default(&o);
    if (opts) o = *opts;
    if (o.refactor_every < 1 || !(o.primal_tol > 0) || !(o.dual_tol > 0))
        return mpdhg::fail(MPDHG_E_ARG, "polish: refactor_every >= 1 and positive tolerances required");
    try {
        const double t0 = mpdhg::now();
        auto solve = [&](const int8_t *cb, const int8_t *rb, const double *xs) {
            mpdhg::Simplex s(*lp, o);
            s.start(cb, rb, xs);
            s.run(t0);
            s.results(*res);
        };
        try {
            solve(col_basis, row_basis, x_start);
        } catch (const mpdhg::FactorError &) {
            // a start basis that stays singular after repair: the slack basis,
            // as the reference does for a singular warm basis
            solve(nullptr, nullptr, nullptr);
        }
        res->time_sec = mpdhg::now() - t0;
    } catch (const std::bad_alloc &) {
        return mpdhg::fail(MPDHG_E_NOMEM, "polish: out of host memory");
    } catch (const std::exception &e) {
        return mpdhg::fail(MPDHG_E_STATE, std::string("polish: ") + e.what());
    }
    return MPDHG_OK;
}

}  // extern "C"
