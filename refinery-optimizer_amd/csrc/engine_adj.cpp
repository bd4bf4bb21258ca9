// Adjoint driver: active-set detection and scaling on the host, batched CGLS
// for A_RB^T mu = w_B on the device (kernels_adj.hip), gradients assembled
// from mu and A^T mu.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "engine.hpp"

namespace mpdhg {

void solve_adjoint(Context &ctx, const mpdhg_lp *lps, int count, mpdhg_adjoint_io *io,
                   double active_tol, double tol, int max_iter) {
    if (!(active_tol > 0.0) || !(tol > 0.0) || max_iter < 1)
        throw Error(MPDHG_E_ARG, "adjoint: active_tol and tol must be > 0, max_iter >= 1");
    std::vector<HostLp> H(static_cast<size_t>(count));
    prepare_all(lps, count, H);
    std::vector<int> act;
    for (int q = 0; q < count; ++q) {
        const HostLp &h = H[static_cast<size_t>(q)];
        mpdhg_adjoint_io &o = io[q];
        if ((h.n > 0 && (!o.x || !o.w)) || (h.m > 0 && !o.y))
            throw Error(MPDHG_E_ARG, "adjoint: x, y and w are required");
        o.basic_cols = o.active_rows = o.iterations = 0;
        o.residual = 0.0;
        if (o.dl_dc) std::fill(o.dl_dc, o.dl_dc + h.n, 0.0);
        if (o.dl_db) std::fill(o.dl_db, o.dl_db + h.m, 0.0);
        if (o.dl_dl) std::fill(o.dl_dl, o.dl_dl + h.n, 0.0);
        if (o.dl_du) std::fill(o.dl_du, o.dl_du + h.n, 0.0);
        if (h.m > 0 && h.n > 0) act.push_back(q);
    }
    const int na = static_cast<int>(act.size());
    if (na == 0) return;
    const BatchLayout L = make_layout(H, act);
    const size_t R = static_cast<size_t>(L.rows), C = static_cast<size_t>(L.cols),
                 Z = static_cast<size_t>(L.nnz);
    std::vector<int> cptr(C + 1), cidx(Z), rptr(R + 1), ridx(Z);
    std::vector<double> cval(Z), rval(Z), dc(C), dr(R), mb(C), mr(R), r0(C), rt0(C);
    std::vector<AdjInst> inst(static_cast<size_t>(na));
    parallel_for_each(na, [&](int a) {
        const int q = act[static_cast<size_t>(a)];
        const HostLp &h = H[static_cast<size_t>(q)];
        const mpdhg_lp &lp = lps[q];
        const mpdhg_adjoint_io &o = io[q];
        const int64_t ro = L.row_off[static_cast<size_t>(a)], co = L.col_off[static_cast<size_t>(a)],
                      zo = L.nnz_off[static_cast<size_t>(a)];
        double *pdr = dr.data() + ro, *pdc = dc.data() + co;
        scale_instance(lp, h, 10, pdr, pdc);
        for (int j = 0; j < h.n; ++j) cptr[static_cast<size_t>(co + j)] = static_cast<int>(zo + h.cptr[j]);
        for (int i = 0; i < h.m; ++i) rptr[static_cast<size_t>(ro + i)] = static_cast<int>(zo + h.rptr[i]);
        for (int j = 0; j < h.n; ++j)
            for (int k = h.cptr[j]; k < h.cptr[j + 1]; ++k) {
                const int d = h.csc2csr[static_cast<size_t>(k)];
                cval[static_cast<size_t>(zo + k)] = lp.vals[k];
                cidx[static_cast<size_t>(zo + k)] = static_cast<int>(ro + h.cidx[static_cast<size_t>(k)]);
                rval[static_cast<size_t>(zo + d)] = lp.vals[k];
                ridx[static_cast<size_t>(zo + d)] = static_cast<int>(co + j);
            }
        // active set: basic columns strictly inside, rows with nonzero duals
        double ymax = 0.0;
        for (int i = 0; i < h.m; ++i) ymax = std::max(ymax, std::fabs(o.y[i]));
        int nb = 0, nr = 0;
        for (int j = 0; j < h.n; ++j) {
            const double xj = o.x[j], marg = active_tol * (1.0 + std::fabs(xj));
            const bool basic = (xj - lp.col_lower[j] > marg) && (lp.col_upper[j] - xj > marg);
            const size_t g = static_cast<size_t>(co + j);
            mb[g] = basic ? 1.0 : 0.0;
            nb += basic;
            r0[g] = mb[g] * pdc[j] * o.w[j];  // scaled right-hand side D_c w_B
            rt0[g] = mb[g] * pdc[j] * r0[g];
        }
        // Active rows: nonzero duals, and rows tight at x.  At a degenerate
        // vertex a tight row can carry a zero dual; leaving it out makes
        // A_RB lose column rank, and A_RB^T mu = w_B inconsistent.
        std::vector<double> ax(static_cast<size_t>(h.m), 0.0);
        for (int j = 0; j < h.n; ++j)
            for (int k = h.cptr[static_cast<size_t>(j)]; k < h.cptr[static_cast<size_t>(j) + 1]; ++k)
                ax[static_cast<size_t>(h.cidx[static_cast<size_t>(k)])] += lp.vals[k] * o.x[j];
        double bnorm = 0.0;
        for (int i = 0; i < h.m; ++i) {
            if (std::isfinite(lp.row_lower[i])) bnorm = std::max(bnorm, std::fabs(lp.row_lower[i]));
            if (std::isfinite(lp.row_upper[i])) bnorm = std::max(bnorm, std::fabs(lp.row_upper[i]));
        }
        const double tight = active_tol * (1.0 + bnorm);
        for (int i = 0; i < h.m; ++i) {
            const double a = ax[static_cast<size_t>(i)];
            const bool active = std::fabs(o.y[i]) > active_tol * (1.0 + ymax) ||
                                (std::isfinite(lp.row_lower[i]) && std::fabs(a - lp.row_lower[i]) <= tight) ||
                                (std::isfinite(lp.row_upper[i]) && std::fabs(a - lp.row_upper[i]) <= tight);
            mr[static_cast<size_t>(ro + i)] = active ? 1.0 : 0.0;
            nr += active;
        }
        AdjInst &s = inst[static_cast<size_t>(a)];
        std::memset(&s, 0, sizeof(s));
        io[q].basic_cols = nb;
        io[q].active_rows = nr;
    });
    cptr[C] = static_cast<int>(Z);
    rptr[R] = static_cast<int>(Z);
    for (int a = 0; a < na; ++a) {
        const HostLp &h = H[static_cast<size_t>(act[static_cast<size_t>(a)])];
        AdjInst &s = inst[static_cast<size_t>(a)];
        const size_t ua = static_cast<size_t>(a);
        s.rgrp_begin = static_cast<int32_t>(L.rgrp_off[ua]);
        s.rgrp_end = static_cast<int32_t>(L.rgrp_off[ua] + (h.m + kGroup - 1) / kGroup);
        s.cgrp_begin = static_cast<int32_t>(L.cgrp_off[ua]);
        s.cgrp_end = static_cast<int32_t>(L.cgrp_off[ua] + (h.n + kGroup - 1) / kGroup);
    }
    hipStream_t st = ctx.stream;
    DBuf<int> d_cptr, d_cidx_s, d_rptr, d_ridx_s;
    DBuf<double> d_cval_s, d_rval_s, d_dc, d_dr, d_mb, d_mr, d_r, d_rt;
    DBuf<double> d_q(C), d_s(R), d_mu(R), d_p(R), d_pt(R);
    DBuf<AdjInst> d_inst;
    d_cptr.upload(cptr, st); d_cidx_s.upload(cidx, st); d_rptr.upload(rptr, st); d_ridx_s.upload(ridx, st);
    d_cval_s.upload(cval, st); d_rval_s.upload(rval, st); d_dc.upload(dc, st); d_dr.upload(dr, st);
    d_mb.upload(mb, st); d_mr.upload(mr, st); d_r.upload(r0, st); d_rt.upload(rt0, st);
    d_inst.upload(inst, st);
    // tile schedules by rule (the CGLS kernels stream FP64 values)
    const DirPlan RP = dir_plan(H, act, L, true), CP = dir_plan(H, act, L, false);
    DevDir RD, CD;
    build_dir(RP, choose_schedule(L.nnz, L.rows, 0, 0), Layout::Auto, DirSource{d_ridx_s.get(), nullptr, d_rval_s.get()},
              false, true, st, RD);
    build_dir(CP, choose_schedule(L.nnz, L.cols, 0, 0), Layout::Auto, DirSource{d_cidx_s.get(), nullptr, d_cval_s.get()},
              false, true, st, CD);
    for (double *p : {d_s.get(), d_mu.get(), d_p.get(), d_pt.get()})
        MPDHG_HIP(hipMemsetAsync(p, 0, R * sizeof(double), st));
    MPDHG_HIP(hipMemsetAsync(d_q.get(), 0, C * sizeof(double), st));
    // per-group partials of |q|^2 (columns) and |s|^2 (rows)
    DBuf<double> d_gq(static_cast<size_t>(L.cgroups)), d_gs(static_cast<size_t>(L.rgroups));
    const TileSet rts = RD.set(), cts = CD.set();
    launch::AdjColArgs CA{d_cptr.get(), CD.idx, CD.v64, d_pt.get(), d_mb.get(), d_dc.get(),
                          d_q.get(), d_r.get(), d_rt.get(), d_gq.get()};
    launch::AdjRowArgs RA{d_rptr.get(), RD.idx, RD.v64, d_rt.get(), d_mr.get(), d_dr.get(),
                          d_s.get(), d_mu.get(), d_p.get(), d_pt.get(), d_gs.get()};
    launch::adj_start(st, rts, d_inst.get(), na, RA, tol, max_iter);
    std::vector<AdjInst> hs(static_cast<size_t>(na));
    // Blocks of 32 CGLS steps, recorded once as a HIP graph and replayed; the
    // host reads each block's done flags while the next block runs (steps of
    // finished instances are no-ops).
    constexpr int kSteps = 32;
    struct StepGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        ~StepGraph() {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
        }
    } graph;
    MPDHG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        for (int u = 0; u < kSteps; ++u) launch::adj_step(st, rts, cts, d_inst.get(), na, CA, RA, tol, max_iter);
    } catch (...) {
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(st, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
    }
    MPDHG_HIP(hipStreamEndCapture(st, &graph.g));
    MPDHG_HIP(hipGraphInstantiate(&graph.e, graph.g, nullptr, nullptr, 0));
    Pinned snap(2 * hs.size() * sizeof(AdjInst));
    AdjInst *const snaps[2] = {static_cast<AdjInst *>(snap.p), static_cast<AdjInst *>(snap.p) + na};
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() { for (auto v : e) if (v) (void)hipEventDestroy(v); }
    } ev;
    for (auto &v : ev.e) MPDHG_HIP(hipEventCreateWithFlags(&v, hipEventDisableTiming));
    for (int k = 0, blk = 0; k < max_iter; k += kSteps, ++blk) {
        MPDHG_HIP(hipGraphLaunch(graph.e, st));
        const int cur = blk & 1;
        MPDHG_HIP(hipMemcpyAsync(snaps[cur], d_inst.get(), hs.size() * sizeof(AdjInst), hipMemcpyDeviceToHost, st));
        MPDHG_HIP(hipEventRecord(ev.e[cur], st));
        if (blk > 0) {
            MPDHG_HIP(hipEventSynchronize(ev.e[cur ^ 1]));
            const AdjInst *sv = snaps[cur ^ 1];
            if (std::all_of(sv, sv + na, [](const AdjInst &s) { return s.done != 0; })) break;
        }
    }
    launch::adj_atmu(st, rts, cts, CA, RA);  // pt = D_r mu, q = A^T pt
    MPDHG_HIP(hipGetLastError());
    std::vector<double> mu, atmu;
    d_pt.download(mu, R, st);
    d_q.download(atmu, C, st);
    MPDHG_HIP(hipMemcpyAsync(hs.data(), d_inst.get(), hs.size() * sizeof(AdjInst), hipMemcpyDeviceToHost, st));
    MPDHG_HIP(hipStreamSynchronize(st));
    for (int a = 0; a < na; ++a) {
        const int q = act[static_cast<size_t>(a)];
        const HostLp &h = H[static_cast<size_t>(q)];
        const mpdhg_lp &lp = lps[q];
        mpdhg_adjoint_io &o = io[q];
        const AdjInst &s = hs[static_cast<size_t>(a)];
        const size_t ro = static_cast<size_t>(L.row_off[static_cast<size_t>(a)]),
                     co = static_cast<size_t>(L.col_off[static_cast<size_t>(a)]);
        o.iterations = s.iters;
        o.residual = s.gamma0 > 0.0 ? std::sqrt(s.gamma / s.gamma0) : 0.0;
        if (o.dl_db)
            for (int i = 0; i < h.m; ++i) o.dl_db[i] = mr[ro + static_cast<size_t>(i)] * mu[ro + static_cast<size_t>(i)];
        for (int j = 0; j < h.n; ++j) {
            if (mb[co + static_cast<size_t>(j)] != 0.0) continue;
            const double g = o.w[j] - atmu[co + static_cast<size_t>(j)];
            const double dl = std::fabs(o.x[j] - lp.col_lower[j]), du = std::fabs(lp.col_upper[j] - o.x[j]);
            if (dl <= du) {
                if (o.dl_dl) o.dl_dl[j] = g;
            } else if (o.dl_du) {
                o.dl_du[j] = g;
            }
        }
    }
}

}  // namespace mpdhg
