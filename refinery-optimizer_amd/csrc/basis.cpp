// Basis identification of GpuSolver::solve_hybrid (gpu_solver.cpp:319-395):
// the hand-off from an approximate PDHG point to an FP64 simplex warm start.
#include <algorithm>
#include <cmath>
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
