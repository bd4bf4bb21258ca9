// Basis identification of GpuSolver::solve_hybrid (gpu_solver.cpp:319-395):
// the hand-off from an approximate PDHG point to an FP64 simplex warm start.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "engine.hpp"
#include "mpdhg.h"
#include "mpdhg_internal.hpp"

extern "C" int mpdhg_identify_basis(const mpdhg_lp *lp, const double *x, int8_t *col_basis,
                                    int8_t *row_basis) {
    if (!lp || !x || (lp->n > 0 && !col_basis) || (lp->m > 0 && !row_basis))
        return mpdhg::fail(MPDHG_E_ARG, "identify_basis: NULL argument");
    try {
        mpdhg::HostLp h;
        mpdhg::prepare_structure(*lp, h);
        const int m = h.m, n = h.n;
        // :320-330 row activities, skipping |x_j| <= 1e-12
        std::vector<double> ax(static_cast<size_t>(m), 0.0);
        for (int j = 0; j < n; ++j) {
            const double xc = x[j];
            if (std::fabs(xc) > 1e-12)
                for (int k = h.cptr[j]; k < h.cptr[j + 1]; ++k)
                    ax[static_cast<size_t>(h.cidx[k])] += lp->vals[k] * xc;
        }
        // :342-368 distance to the nearer finite bound (1e30 when infinite)
        const size_t N = static_cast<size_t>(n) + static_cast<size_t>(m);
        std::vector<double> dist(N);
        std::vector<int8_t> nb(N);
        for (size_t t = 0; t < N; ++t) {
            double v, lo, hi;
            if (t < static_cast<size_t>(n)) {
                v = x[t];
                lo = lp->col_lower[t];
                hi = lp->col_upper[t];
            } else {
                const size_t i = t - static_cast<size_t>(n);
                v = ax[i];
                lo = lp->row_lower[i];
                hi = lp->row_upper[i];
            }
            const double dl = (lo > -1e15) ? std::max(0.0, v - lo) : 1e30;
            const double du = (hi < 1e15) ? std::max(0.0, hi - v) : 1e30;
            dist[t] = std::min(dl, du);
            nb[t] = (dl <= du) ? MPDHG_BASIS_AT_LOWER : MPDHG_BASIS_AT_UPPER;
        }
        // :371-373 most interior first; equal distances keep index order
        std::vector<size_t> order(N);
        std::iota(order.begin(), order.end(), size_t{0});
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return dist[a] > dist[b]; });
        // :375-395 top m Basic, the rest at their nearer bound
        for (size_t k = 0; k < N; ++k) {
            const size_t id = order[k];
            const int8_t s = (k < static_cast<size_t>(m)) ? int8_t(MPDHG_BASIS_BASIC) : nb[id];
            if (id < static_cast<size_t>(n)) col_basis[id] = s;
            else row_basis[id - static_cast<size_t>(n)] = s;
        }
        return MPDHG_OK;
    } catch (const mpdhg::Error &e) {
        return mpdhg::fail(e.code, e.what());
    } catch (const std::exception &e) {
        return mpdhg::fail(MPDHG_E_NOMEM, e.what());
    }
}

// GpuSolver::solve_hybrid (gpu_solver.cpp:309-417): PDHG on the device, the
// basis hand-off, the FP64 polish from that basis and, when the polish does
// not end optimal, the same simplex from the slack basis (:413-416).
extern "C" int mpdhg_solve_hybrid(mpdhg_ctx *ctx, const mpdhg_lp *lp, const mpdhg_config *cfg,
                                  const mpdhg_polish_opts *popts, mpdhg_result *pdhg,
                                  mpdhg_polish_result *out) {
    if (!lp || !out) return mpdhg::fail(MPDHG_E_ARG, "solve_hybrid: NULL lp or result");
    if (lp->m < 0 || lp->n < 0) return mpdhg::fail(MPDHG_E_ARG, "solve_hybrid: malformed LP");
    const auto t0 = std::chrono::steady_clock::now();
    mpdhg_result local{};
    mpdhg_result *pr = pdhg ? pdhg : &local;
    double *const user_x = pr->x;
    std::vector<double> xbuf;
    try {
        if (!user_x) {
            xbuf.resize(static_cast<size_t>(lp->n));
            pr->x = xbuf.data();
        }
        std::vector<int8_t> cb(static_cast<size_t>(lp->n)), rb(static_cast<size_t>(lp->m));
        int rc = mpdhg_solve(ctx, lp, cfg, pr);
        if (rc == MPDHG_OK) rc = mpdhg_identify_basis(lp, pr->x, cb.data(), rb.data());
        const std::vector<double> xs(pr->x, pr->x + lp->n);
        pr->x = user_x;
        if (rc != MPDHG_OK) return rc;
        rc = mpdhg_polish(lp, cb.data(), rb.data(), xs.data(), popts, out);
        if (rc == MPDHG_OK && out->status != MPDHG_STATUS_OPTIMAL) {
            const int64_t first = out->iterations;
            rc = mpdhg_polish(lp, nullptr, nullptr, nullptr, popts, out);
            out->iterations += first;
        }
        if (rc != MPDHG_OK) return rc;
    } catch (const std::bad_alloc &) {
        pr->x = user_x;
        return mpdhg::fail(MPDHG_E_NOMEM, "solve_hybrid: out of host memory");
    } catch (const std::exception &e) {
        pr->x = user_x;
        return mpdhg::fail(MPDHG_E_STATE, std::string("solve_hybrid: ") + e.what());
    }
    out->time_sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return MPDHG_OK;
}

namespace mpdhg {

// The error of each PDHG result against an FP64 optimal vertex: the
// crossover of solve_hybrid (basis identification, then the polish from that
// basis and the PDHG point, then from the slack basis if that does not end
// optimal, gpu_solver.cpp:397-416), on host threads, for the LPs with at most
// max_rows rows.
void vertex_check(const mpdhg_lp *lps, int count, const std::vector<const double *> &x, int max_rows,
                  mpdhg_result *res) {
    parallel_for_each(count, [&](int q) {
        mpdhg_result &r = res[q];
        r.vertex_status = MPDHG_STATUS_UNKNOWN;
        r.vertex_objective = std::numeric_limits<double>::quiet_NaN();
        r.vertex_rel_error = std::numeric_limits<double>::quiet_NaN();
        const mpdhg_lp &lp = lps[q];
        const double *xq = x[static_cast<size_t>(q)];
        if (!xq || lp.m <= 0 || lp.n <= 0 || lp.m > max_rows) return;
        std::vector<int8_t> cb(static_cast<size_t>(lp.n)), rb(static_cast<size_t>(lp.m));
        if (mpdhg_identify_basis(&lp, xq, cb.data(), rb.data()) != MPDHG_OK) return;
        mpdhg_polish_opts o;
        mpdhg_polish_opts_default(&o);
        o.time_limit_sec = 60.0;
        mpdhg_polish_result pr;
        std::memset(&pr, 0, sizeof(pr));
        if (mpdhg_polish(&lp, cb.data(), rb.data(), xq, &o, &pr) != MPDHG_OK) return;
        if (pr.status != MPDHG_STATUS_OPTIMAL && pr.status != MPDHG_STATUS_TIME_LIMIT) {
            std::memset(&pr, 0, sizeof(pr));
            if (mpdhg_polish(&lp, nullptr, nullptr, nullptr, &o, &pr) != MPDHG_OK) return;
        }
        r.vertex_status = pr.status;
        r.vertex_objective = pr.objective;
        if (pr.status == MPDHG_STATUS_OPTIMAL)
            r.vertex_rel_error = std::fabs(r.primal_objective - pr.objective) / std::max(1.0, std::fabs(pr.objective));
    });
}

}  // namespace mpdhg

