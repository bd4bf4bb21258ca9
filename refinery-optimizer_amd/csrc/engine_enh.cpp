// Enhanced-mode driver: per-instance scaling on the host, then a fully
// device-resident loop.  Each check cycle is F fused primal/dual launch
// pairs followed by the FP64 check launches; restart, primal-weight and
// refinement decisions are taken on the device by the per-instance finalize
// kernel, so the host only polls a status word per instance, one cycle
// behind the device.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <limits>
#include <mutex>
#include <thread>

#include "engine.hpp"

namespace mpdhg {

namespace {

// Stream bytes per iteration above which the fused kernels load the matrix
// stream non-temporally: the size of MI355X's last-level (Infinity) cache.
constexpr double kNtStreamBytes = 256.0 * 1024 * 1024;

// The batch's host arrays, written in full by the per-instance preparation
// (hence uninitialised: see HostArray), except the optional warm-start duals.
struct EnhHost {
    HostArray<int> cptr, cidx, rptr, perm;
    HostArray<float> cs, ls, us, rls, rus, x0, powv;
    HostArray<double> cval64, c64, l64, u64, dc, rl64, ru64, dr;
    std::vector<float> y0;
    std::vector<EnhInst> inst;
};

double finite_or_zero(float a, float b) {
    return std::isfinite(a) ? static_cast<double>(a) : (std::isfinite(b) ? static_cast<double>(b) : 0.0);
}

}  // namespace

void scale_instance(const mpdhg_lp &lp, const HostLp &h, int ruiz, double *dr, double *dc) {
    const int m = h.m, n = h.n;
    std::vector<double> rs(static_cast<size_t>(m)), cs(static_cast<size_t>(n));
    for (int i = 0; i < m; ++i) dr[i] = 1.0;
    for (int j = 0; j < n; ++j) dc[j] = 1.0;
    for (int pass = 0; pass <= ruiz; ++pass) {
        const bool l1 = (pass == ruiz);
        std::fill(rs.begin(), rs.end(), 0.0);
        for (int j = 0; j < n; ++j) {
            double acc = 0.0;
            for (int k = h.cptr[j]; k < h.cptr[j + 1]; ++k) {
                const int i = h.cidx[static_cast<size_t>(k)];
                const double a = std::fabs(lp.vals[k]) * dr[i] * dc[j];
                if (l1) {
                    acc += a;
                    rs[static_cast<size_t>(i)] += a;
                } else {
                    if (a > acc) acc = a;
                    if (a > rs[static_cast<size_t>(i)]) rs[static_cast<size_t>(i)] = a;
                }
            }
            cs[static_cast<size_t>(j)] = acc;
        }
        for (int i = 0; i < m; ++i)
            if (rs[static_cast<size_t>(i)] > 0.0) dr[i] /= std::sqrt(rs[static_cast<size_t>(i)]);
        for (int j = 0; j < n; ++j)
            if (cs[static_cast<size_t>(j)] > 0.0) dc[j] /= std::sqrt(cs[static_cast<size_t>(j)]);
    }
}

namespace {

// f(i) for i in [0, count) on up to hardware_concurrency threads.  An
// exception thrown by f stops the remaining work and is rethrown here, after
// every thread has joined.
template <typename F>
void parallel_for(int count, F &&f) {
    const int nt = std::max(1, std::min<int>(count, static_cast<int>(std::thread::hardware_concurrency())));
    if (nt <= 1) {
        for (int i = 0; i < count; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<int> next{0};
    std::mutex mu;
    std::exception_ptr first;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&] {
            try {
                for (int i = next++; i < count; i = next++) f(i);
            } catch (...) {
                std::lock_guard<std::mutex> lock(mu);
                if (!first) first = std::current_exception();
                next = count;
            }
        });
    for (auto &th : pool) th.join();
    if (first) std::rethrow_exception(first);
}

}  // namespace

void parallel_for_each(int count, const std::function<void(int)> &f) { parallel_for(count, f); }

void solve_enhanced(Context &ctx, const mpdhg_lp *lps, int count, const mpdhg_config &cfg,
                    const mpdhg_warm_start *warm, mpdhg_result *res) {
    const double t_call = now_sec();
    const int F = cfg.check_frequency, K = cfg.max_iterations;
    const double tol = cfg.tol_kkt > 0.0 ? cfg.tol_kkt
                                         : std::max<double>(cfg.tol_primal, std::max(cfg.tol_dual, cfg.tol_gap));
    if (cfg.reflection < 0.0f || cfg.reflection > 1.0f) throw Error(MPDHG_E_ARG, "reflection must be in [0, 1]");
    if (cfg.ruiz_iterations < 0) throw Error(MPDHG_E_ARG, "ruiz_iterations must be >= 0");
    // MPDHG_SETUP_TRACE set: the set-up's phases are timed (each after a
    // stream synchronisation) and printed to stderr.
    const bool trace = std::getenv("MPDHG_SETUP_TRACE") != nullptr;
    double t_mark = t_call;
    auto mark = [&](const char *what) {
        if (!trace) return;
        MPDHG_HIP(hipStreamSynchronize(ctx.stream));
        const double t = now_sec();
        std::fprintf(stderr, "[mpdhg setup] %-34s %8.2f ms\n", what, 1e3 * (t - t_mark));
        t_mark = t;
    };
    std::vector<HostLp> H(static_cast<size_t>(count));
    prepare_all(lps, count, H);
    std::vector<int> act;
    for (int q = 0; q < count; ++q) {
        HostLp &h = H[static_cast<size_t>(q)];
        mpdhg_result &r = res[q];
        if (h.m == 0 || h.n == 0) {
            r.status = MPDHG_STATUS_OPTIMAL;
        } else if (K <= 0) {  // no iteration: report the projected origin
            std::vector<double> x(static_cast<size_t>(h.n)), y(static_cast<size_t>(h.m), 0.0);
            for (int j = 0; j < h.n; ++j)
                x[static_cast<size_t>(j)] = std::max(lps[q].col_lower[j], std::min(lps[q].col_upper[j], 0.0));
            host_kkt(lps[q], h, x, y, r);
            double obj = 0;
            for (int j = 0; j < h.n; ++j) obj += lps[q].c[j] * x[static_cast<size_t>(j)];
            r.primal_objective = obj;
            if (r.x) std::memcpy(r.x, x.data(), x.size() * sizeof(double));
            if (r.y) std::fill(r.y, r.y + h.m, 0.0);
            r.status = MPDHG_STATUS_ITERATION_LIMIT;
        } else {
            act.push_back(q);
        }
    }
    stratify_by_size(act, H);
    const int na = static_cast<int>(act.size());
    if (na == 0) {
        for (int q = 0; q < count; ++q) res[q].time_wall_sec = now_sec() - t_call;
        return;
    }
    const BatchLayout L = make_layout(H, act);
    const size_t R = static_cast<size_t>(L.rows), C = static_cast<size_t>(L.cols),
                 Z = static_cast<size_t>(L.nnz);
    EnhHost hh;
    hh.cptr.resize(C + 1); hh.cidx.resize(Z); hh.rptr.resize(R + 1); hh.perm.resize(Z); hh.cval64.resize(Z);
    hh.cs.resize(C); hh.ls.resize(C); hh.us.resize(C); hh.x0.resize(C); hh.powv.resize(C);
    hh.c64.resize(C); hh.l64.resize(C); hh.u64.resize(C); hh.dc.resize(C);
    hh.rls.resize(R); hh.rus.resize(R); hh.rl64.resize(R); hh.ru64.resize(R); hh.dr.resize(R);
    hh.y0.assign(R, 0.0f);
    hh.inst.resize(static_cast<size_t>(na));
    std::vector<double> pow0(static_cast<size_t>(na));
    mark("CSR builds, layout, host arrays");

    // ---- per-instance host preparation, part 1 (independent, in parallel):
    // the batch's structure and its FP64 values in CSC order; an LP of more
    // than kParallelCsrNnz nonzeros is cut into parts of columns and rows ----
    std::vector<std::pair<int, int>> parts;  // (instance, part of instance_parts)
    auto instance_parts = [&](int a) {
        return 1 + H[static_cast<size_t>(act[static_cast<size_t>(a)])].nnz / kParallelCsrNnz;
    };
    for (int a = 0; a < na; ++a)
        for (int t = 0; t < instance_parts(a); ++t) parts.emplace_back(a, t);
    parallel_for(static_cast<int>(parts.size()), [&](int w) {
        const int a = parts[static_cast<size_t>(w)].first, t = parts[static_cast<size_t>(w)].second;
        const int np = instance_parts(a);
        const int q = act[static_cast<size_t>(a)];
        const HostLp &h = H[static_cast<size_t>(q)];
        const mpdhg_lp &lp = lps[q];
        const int64_t ro = L.row_off[static_cast<size_t>(a)], co = L.col_off[static_cast<size_t>(a)],
                      zo = L.nnz_off[static_cast<size_t>(a)];
        const int j0 = static_cast<int>(int64_t(h.n) * t / np), j1 = static_cast<int>(int64_t(h.n) * (t + 1) / np);
        const int i0 = static_cast<int>(int64_t(h.m) * t / np), i1 = static_cast<int>(int64_t(h.m) * (t + 1) / np);
        for (int j = j0; j < j1; ++j) hh.cptr[static_cast<size_t>(co + j)] = static_cast<int>(zo + h.cptr[j]);
        for (int i = i0; i < i1; ++i) hh.rptr[static_cast<size_t>(ro + i)] = static_cast<int>(zo + h.rptr[i]);
        for (int k = h.cptr[static_cast<size_t>(j0)]; k < h.cptr[static_cast<size_t>(j1)]; ++k) {
            const size_t gk = static_cast<size_t>(zo + k);
            hh.cidx[gk] = static_cast<int>(ro + h.cidx[static_cast<size_t>(k)]);
            hh.cval64[gk] = lp.vals[k];
            hh.perm[gk] = static_cast<int>(zo + h.csc2csr[static_cast<size_t>(k)]);
        }
    });
    hh.cptr[C] = static_cast<int>(Z);
    hh.rptr[R] = static_cast<int>(Z);
    mark("structure and FP64 values");

    // ---- on the device: the CSR copy, Ruiz / Pock-Chambolle scaling and the
    // scaled FP32 values (the matrix in storage order: the layout source) ----
    hipStream_t st = ctx.stream;
    DBuf<int> d_cptr, d_cidx_s, d_rptr, d_ridx_s(Z);
    DBuf<float> d_cval_s(Z), d_rval_s(Z), d_cs, d_ls, d_us, d_rls, d_rus;
    DBuf<double> d_cval64_s, d_rval64_s(Z), d_c64, d_l64, d_u64, d_dc(C), d_rl64, d_ru64, d_dr(R);
    {
        DBuf<int> d_perm;
        DBuf<double> d_rs(R), d_csum(C);
        d_cptr.upload(hh.cptr.data(), hh.cptr.size(), st);
        d_cidx_s.upload(hh.cidx.data(), hh.cidx.size(), st);
        d_rptr.upload(hh.rptr.data(), hh.rptr.size(), st);
        d_perm.upload(hh.perm.data(), hh.perm.size(), st);
        d_cval64_s.upload(hh.cval64.data(), hh.cval64.size(), st);
        mark("structure and value uploads");
        launch::scale_batch(st, static_cast<int>(R), static_cast<int>(C), d_cptr.get(), d_cidx_s.get(), d_rptr.get(),
                            d_perm.get(), d_cval64_s.get(), d_rval64_s.get(), d_ridx_s.get(), cfg.ruiz_iterations,
                            d_dr.get(), d_dc.get(), d_rs.get(), d_csum.get(), d_cval_s.get(), d_rval_s.get());
        MPDHG_HIP(hipGetLastError());
        MPDHG_HIP(hipMemcpyAsync(hh.dr.data(), d_dr.get(), R * sizeof(double), hipMemcpyDeviceToHost, st));
        MPDHG_HIP(hipMemcpyAsync(hh.dc.data(), d_dc.get(), C * sizeof(double), hipMemcpyDeviceToHost, st));
        MPDHG_HIP(hipStreamSynchronize(st));
    }
    mark("scaling on the device");

    // ---- per-instance host preparation, part 2: the scaled vectors, norms
    // and the instance state (D_r, D_c from the device) ----
    parallel_for(na, [&](int a) {
        const int q = act[static_cast<size_t>(a)];
        const HostLp &h = H[static_cast<size_t>(q)];
        const mpdhg_lp &lp = lps[q];
        const int64_t ro = L.row_off[static_cast<size_t>(a)], co = L.col_off[static_cast<size_t>(a)];
        const double *dr = hh.dr.data() + ro, *dc = hh.dc.data() + co;
        double cn2 = 0.0, bn2 = 0.0, cinf = 0.0, binf = 0.0, pn = 0.0;
        // a warm start in the scaled space: x^ = x / D_c, y^ = y / D_r
        const mpdhg_warm_start *ws = warm ? &warm[q] : nullptr;
        for (int j = 0; j < h.n; ++j) {
            const size_t g = static_cast<size_t>(co + j);
            hh.c64[g] = lp.c[j];
            hh.l64[g] = lp.col_lower[j];
            hh.u64[g] = lp.col_upper[j];
            hh.cs[g] = static_cast<float>(lp.c[j] * dc[j]);
            hh.ls[g] = static_cast<float>(lp.col_lower[j] / dc[j]);
            hh.us[g] = static_cast<float>(lp.col_upper[j] / dc[j]);
            float v = (ws && ws->x) ? static_cast<float>(ws->x[j] / dc[j]) : 0.0f;
            if (v < hh.ls[g]) v = hh.ls[g];
            if (v > hh.us[g]) v = hh.us[g];
            hh.x0[g] = v;
            cn2 += static_cast<double>(hh.cs[g]) * hh.cs[g];
            cinf = std::max(cinf, std::fabs(lp.c[j]));
            const float pv = 1.0f + 0.5f * static_cast<float>((static_cast<uint32_t>(j) * 2654435761u) % 1000u) / 1000.0f;
            hh.powv[g] = pv;
            pn += static_cast<double>(pv) * pv;
        }
        if (ws && ws->y) {
            const double sg = cfg.dual_sign ? -1.0 : 1.0;
            for (int i = 0; i < h.m; ++i)
                hh.y0[static_cast<size_t>(ro + i)] = static_cast<float>(sg * ws->y[i] / dr[i]);
        }
        for (int i = 0; i < h.m; ++i) {
            const size_t g = static_cast<size_t>(ro + i);
            hh.rl64[g] = lp.row_lower[i];
            hh.ru64[g] = lp.row_upper[i];
            hh.rls[g] = static_cast<float>(lp.row_lower[i] * dr[i]);
            hh.rus[g] = static_cast<float>(lp.row_upper[i] * dr[i]);
            const double b = finite_or_zero(hh.rus[g], hh.rls[g]);
            bn2 += b * b;
            if (std::isfinite(lp.row_lower[i])) binf = std::max(binf, std::fabs(lp.row_lower[i]));
            if (std::isfinite(lp.row_upper[i])) binf = std::max(binf, std::fabs(lp.row_upper[i]));
        }
        const double cn = std::sqrt(cn2), bn = std::sqrt(bn2);
        EnhInst &s = hh.inst[static_cast<size_t>(a)];
        std::memset(&s, 0, sizeof(s));
        s.status = kRunning;
        s.restart_at = -1;
        s.max_iter = K;
        s.omega = (ws && ws->primal_weight > 0.0) ? ws->primal_weight
                                                   : ((cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0);
        if (ws && ws->step_size > 0.0) s.eta = ws->step_size;
        s.r_anchor = s.r_prev = -1.0;
        s.bnorm = binf;
        s.cnorm = cinf;
        s.tol = tol;
        s.alpha_p = s.alpha_d = s.inv_ap = s.inv_ad = 1.0;
        s.best_kkt = HUGE_VAL;
        // Refinement baseline: HUGE_VAL lets the first check set it.  A warm start
        // with its primal weight is a previous solve's point, already at the FP32
        // iterates' precision; the largest finite double keeps the first check from
        // setting the baseline, so the first restart re-centres it in FP64 and
        // the iterates carry only the correction.
        s.round_start_kkt = (ws && ws->primal_weight > 0.0) ? std::numeric_limits<double>::max() : HUGE_VAL;
        s.pow_norm2 = pn;
        pow0[static_cast<size_t>(a)] = pn;
    });
    for (int a = 0; a < na; ++a) {
        const int q = act[static_cast<size_t>(a)];
        const HostLp &h = H[static_cast<size_t>(q)];
        EnhInst &s = hh.inst[static_cast<size_t>(a)];
        const size_t ua = static_cast<size_t>(a);
        s.rgrp_begin = static_cast<int32_t>(L.rgrp_off[ua]);
        s.rgrp_end = static_cast<int32_t>(L.rgrp_off[ua] + (h.m + kGroup - 1) / kGroup);
        s.cgrp_begin = static_cast<int32_t>(L.cgrp_off[ua]);
        s.cgrp_end = static_cast<int32_t>(L.cgrp_off[ua] + (h.n + kGroup - 1) / kGroup);
    }

    mark("scaled vectors, norms");
    // ---- device buffers: the vectors ----
    d_cs.upload(hh.cs.data(), hh.cs.size(), st); d_ls.upload(hh.ls.data(), hh.ls.size(), st); d_us.upload(hh.us.data(), hh.us.size(), st);
    d_rls.upload(hh.rls.data(), hh.rls.size(), st); d_rus.upload(hh.rus.data(), hh.rus.size(), st);
    d_c64.upload(hh.c64.data(), hh.c64.size(), st); d_l64.upload(hh.l64.data(), hh.l64.size(), st); d_u64.upload(hh.u64.data(), hh.u64.size(), st);
    d_rl64.upload(hh.rl64.data(), hh.rl64.size(), st); d_ru64.upload(hh.ru64.data(), hh.ru64.size(), st);
    DBuf<float> d_x(C), d_x0(C), d_xp(C), d_xbar(C), d_y(R), d_y0(R), d_yp(R), d_ybar(R);
    DBuf<double> d_xbase(C), d_xchk(C), d_lam(C), d_xbest(C), d_lambest(C);
    DBuf<double> d_ybase(R), d_ychk(R), d_ax(R), d_ybest(R), d_axbest(R);
    DBuf<EnhInst> d_inst;
    // Under an arena (the no-allocation entry point) the solve allocates
    // nothing beyond it: no schedule search, no profiling events, no graph,
    // and status polls that wait on the stream instead of on events.
    const bool noalloc = arena_active();
    // per-group partials of the decision sums and maxima (quantity-major)
    const int gcn = static_cast<int>(L.cgroups), grn = static_cast<int>(L.rgroups);
    DBuf<double> d_gc(static_cast<size_t>(kGcCount) * L.cgroups), d_gr(static_cast<size_t>(kGrCount) * L.rgroups);

    mark("vector uploads, buffers");
    // ---- tile schedules: searched on this batch's own matrix (the iterate
    // buffers serve as scratch before they are initialised), or by rule ----
    const Layout lay = static_cast<Layout>(cfg.spmv_layout);
    const DirPlan RP = dir_plan(H, act, L, true), CP = dir_plan(H, act, L, false);
    const DirSource RS{d_ridx_s.get(), d_rval_s.get(), d_rval64_s.get()};
    const DirSource CS{d_cidx_s.get(), d_cval_s.get(), d_cval64_s.get()};
    DevDir RD, CD;
    double tune_ms = 0.0;
    if (tune_wanted(cfg, L.nnz) && !noalloc) {
        std::vector<TuneTrial> trials;
        tune_ms += tune_dir(RP, true, lay, RS, d_rptr.get(), cfg.spmv_block, cfg.spmv_chunk, true, true,
                            d_x.get(), d_y.get(), nullptr, st, RD, trials);
        tune_ms += tune_dir(CP, false, lay, CS, d_cptr.get(), cfg.spmv_block, cfg.spmv_chunk, true, true,
                            d_y.get(), d_x.get(), d_gc.get(), st, CD, trials);
    } else {
        build_dir(RP, choose_schedule(L.nnz, L.rows, cfg.spmv_block, cfg.spmv_chunk), lay, RS, true, true, st, RD);
        build_dir(CP, choose_schedule(L.nnz, L.cols, cfg.spmv_block, cfg.spmv_chunk), lay, CS, true, true, st, CD);
    }
    mark("tile layouts, schedule search");
    // ---- column bands of the A xbar product of the fused dual kernel (one
    // LP whose xbar exceeds an XCD's L2; bit-identical to the unbanded
    // product, which the no-allocation solve keeps) ----
    // (Not for an LP with split rows, kSplitRow: bands would cut their segments.)
    bool split_rows = false;
    if (na == 1) {
        const std::vector<int> &rp = H[static_cast<size_t>(act[0])].rptr;
        for (size_t i = 0; i + 1 < rp.size() && !split_rows; ++i) split_rows = rp[i + 1] - rp[i] > kSplitRow;
    }
    const int bands = (na == 1 && !noalloc && !split_rows) ? band_count(cfg.spmv_bands, L.cols) : 1;
    ColumnBands BD;
    if (bands > 1) {
        // the bands are cut on the host from the CSR the device built
        std::vector<int> ridx;
        std::vector<float> rval;
        d_ridx_s.download(ridx, Z, st);
        d_rval_s.download(rval, Z, st);
        MPDHG_HIP(hipStreamSynchronize(st));
        tune_ms += build_bands(bands, static_cast<int>(R), L.cols, hh.rptr.data(), ridx.data(), rval.data(), cfg, lay,
                               tune_wanted(cfg, L.nnz) && !noalloc, d_x.get(), d_y.get(), st, BD);
    }
    for (int a = 0; a < na; ++a) {
        EnhInst &s = hh.inst[static_cast<size_t>(a)];
        s.row_tile_begin = RD.first[static_cast<size_t>(a)];
        s.row_tile_end = RD.first[static_cast<size_t>(a) + 1];
        s.col_tile_begin = CD.first[static_cast<size_t>(a)];
        s.col_tile_end = CD.first[static_cast<size_t>(a) + 1];
    }
    d_inst.upload(hh.inst, st);
    DBuf<int32_t> d_status(static_cast<size_t>(na));
    MPDHG_HIP(hipMemcpyAsync(d_x.get(), hh.x0.data(), C * sizeof(float), hipMemcpyHostToDevice, st));
    MPDHG_HIP(hipMemcpyAsync(d_x0.get(), hh.x0.data(), C * sizeof(float), hipMemcpyHostToDevice, st));
    MPDHG_HIP(hipMemcpyAsync(d_xp.get(), hh.x0.data(), C * sizeof(float), hipMemcpyHostToDevice, st));
    for (float *p : {d_y.get(), d_y0.get(), d_yp.get()})
        MPDHG_HIP(hipMemcpyAsync(p, hh.y0.data(), R * sizeof(float), hipMemcpyHostToDevice, st));
    MPDHG_HIP(hipMemsetAsync(d_ybar.get(), 0, R * sizeof(float), st));
    for (double *p : {d_xbase.get(), d_xbest.get(), d_lambest.get()})
        MPDHG_HIP(hipMemsetAsync(p, 0, C * sizeof(double), st));
    for (double *p : {d_ybase.get(), d_ybest.get(), d_axbest.get()})
        MPDHG_HIP(hipMemsetAsync(p, 0, R * sizeof(double), st));
    {
        std::vector<int32_t> run(static_cast<size_t>(na), kRunning);
        d_status.upload(run, st);
    }
    TileSet rts = RD.set(), cts = CD.set();
    // Index and value stream of the listed instances, both directions, per
    // iteration: above kNtStreamBytes it cannot stay in the caches between
    // products, so the fused kernels read it non-temporally.
    auto stream_bytes = [&](auto &&listed) {
        double b = 0.0;
        for (int a = 0; a < na; ++a)
            if (listed(a)) {
                const int64_t z1 = (a + 1 < na) ? L.nnz_off[static_cast<size_t>(a) + 1] : L.nnz;
                b += 16.0 * static_cast<double>(z1 - L.nnz_off[static_cast<size_t>(a)]);
            }
        return b;
    };
    rts.stream_nt = cts.stream_nt = stream_bytes([](int) { return true; }) > kNtStreamBytes;
    for (TileSet &bt : BD.sets) bt.stream_nt = rts.stream_nt;
    // power iteration vectors reuse the iterate buffers before the solve
    MPDHG_HIP(hipMemcpyAsync(d_xbar.get(), hh.powv.data(), C * sizeof(float), hipMemcpyHostToDevice, st));
    MPDHG_HIP(hipStreamSynchronize(st));
    mark("bands, instance state, start point");
    const double t_loop = now_sec();

    // ---- split rows (more than kSplitRow nonzeros, either direction): a
    // pre-pass spreads their blocks over many workgroups and leaves the block
    // sums for the fused kernels' owner lanes to add in order (the
    // no-allocation solve forms them inside their tiles instead: same bits) ----
    struct SplitDir {
        std::vector<SplitChunk> host;
        std::vector<int> rowseg;
        DBuf<SplitChunk> d_items;
        DBuf<int> d_rowseg;
        DBuf<float> d_blk;  // block sums
        int n = 0;
    } SR, SC;
    auto plan_split = [&](bool csr, SplitDir &D) {
        int64_t blocks = 0;
        for (int a = 0; a < na; ++a) {
            const HostLp &h = H[static_cast<size_t>(act[static_cast<size_t>(a)])];
            const std::vector<int> &ptr = csr ? h.rptr : h.cptr;
            const int cnt = csr ? h.m : h.n;
            const int64_t roff = csr ? L.row_off[static_cast<size_t>(a)] : L.col_off[static_cast<size_t>(a)];
            const int64_t zoff = L.nnz_off[static_cast<size_t>(a)];
            for (int i = 0; i < cnt; ++i) {
                const int k0 = ptr[static_cast<size_t>(i)], len = ptr[static_cast<size_t>(i) + 1] - k0;
                if (len <= kSplitRow) continue;
                if (D.rowseg.empty()) D.rowseg.assign(csr ? R : C, -1);
                D.rowseg[static_cast<size_t>(roff + i)] = static_cast<int>(blocks);
                for (int c0 = 0; c0 < len; c0 += kSplitBlockNnz)
                    D.host.push_back(SplitChunk{static_cast<int32_t>(zoff + k0 + c0),
                                                static_cast<int32_t>(zoff + k0 + std::min(len, c0 + kSplitBlockNnz)),
                                                static_cast<int32_t>(blocks + c0 / kSplitBlockNnz), a});
                blocks += (len + kSplitBlockNnz - 1) / kSplitBlockNnz;
            }
        }
        D.n = static_cast<int>(D.host.size());
        if (!D.n) return;
        D.d_items.upload(D.host, st);
        D.d_rowseg.upload(D.rowseg, st);
        D.d_blk.resize(static_cast<size_t>(blocks));
    };
    if (!noalloc) {
        plan_split(true, SR);
        plan_split(false, SC);
    }
    // ---- step size: power iteration on (D_r A D_c)^T (D_r A D_c), unless
    // every LP brings its step size with its warm start ----
    bool all_steps = true;
    for (int a = 0; a < na; ++a) all_steps = all_steps && hh.inst[static_cast<size_t>(a)].eta > 0.0;
    for (int p = 0; p < (all_steps ? 0 : 64); ++p) {
        launch::pow_scale(st, cts, d_inst.get(), d_xbar.get(), d_xp.get());
        if (SR.n)
            launch::split_segments(st, SR.d_items.get(), SR.n, d_ridx_s.get(), d_rval_s.get(), d_xp.get(), nullptr,
                                   nullptr, nullptr, 0, SR.d_blk.get());
        launch::pow_rows(st, rts, d_rptr.get(), RD.idx, RD.v32, d_xp.get(), d_yp.get(), SR.n ? SR.d_blk.get() : nullptr,
                         SR.n ? SR.d_rowseg.get() : nullptr);
        if (SC.n)
            launch::split_segments(st, SC.d_items.get(), SC.n, d_cidx_s.get(), d_cval_s.get(), d_yp.get(), nullptr,
                                   nullptr, nullptr, 0, SC.d_blk.get());
        launch::pow_cols(st, cts, d_cptr.get(), CD.idx, CD.v32, d_yp.get(), d_xbar.get(),
                         d_gc.get() + static_cast<size_t>(kGcPow) * L.cgroups, SC.n ? SC.d_blk.get() : nullptr,
                         SC.n ? SC.d_rowseg.get() : nullptr);
        launch::pow_fin(st, d_inst.get(), na, d_gc.get() + static_cast<size_t>(kGcPow) * L.cgroups);
    }
    MPDHG_HIP(hipGetLastError());
    std::vector<EnhInst> hs(static_cast<size_t>(na));
    MPDHG_HIP(hipMemcpyAsync(hs.data(), d_inst.get(), hs.size() * sizeof(EnhInst), hipMemcpyDeviceToHost, st));
    MPDHG_HIP(hipStreamSynchronize(st));
    for (int a = 0; a < na; ++a) {
        EnhInst &s = hh.inst[static_cast<size_t>(a)];
        const double sig = std::sqrt(std::sqrt(hs[static_cast<size_t>(a)].pow_norm2));
        if (!(s.eta > 0.0)) s.eta = 0.998 / (sig > 1e-12 ? sig : 1.0);
        s.tau = static_cast<float>(s.eta / s.omega);
        s.sigma = static_cast<float>(s.eta * s.omega);
        s.inv_sigma = static_cast<float>(1.0 / (s.eta * s.omega));
        s.pow_norm2 = hs[static_cast<size_t>(a)].pow_norm2;
    }
    d_inst.upload(hh.inst, st);
    MPDHG_HIP(hipMemcpyAsync(d_xp.get(), hh.x0.data(), C * sizeof(float), hipMemcpyHostToDevice, st));
    MPDHG_HIP(hipMemcpyAsync(d_yp.get(), hh.y0.data(), R * sizeof(float), hipMemcpyHostToDevice, st));

    // ---- iteration loop ----
    // The first iteration of the current check cycle lives on the device
    // (advanced by the last launch of each cycle), so one recorded cycle
    // replays unchanged for every full cycle.
    DBuf<int> d_itb(1);
    MPDHG_HIP(hipMemsetAsync(d_itb.get(), 0, sizeof(int), st));
    const float rho = cfg.reflection;
    launch::EnhPrimal P{d_cptr.get(), CD.idx, CD.v32, d_y.get(), d_yp.get(), d_cs.get(),
                        d_ls.get(), d_us.get(), d_x.get(), d_x0.get(), d_xp.get(), d_xbar.get(),
                        d_xbase.get(), d_dc.get(), d_xchk.get(), d_gc.get(), gcn, d_itb.get(), rho};
    launch::EnhDual D{d_rptr.get(), RD.idx, RD.v32, d_xbar.get(), d_rls.get(), d_rus.get(),
                      d_ybar.get(), d_y.get(), d_y0.get(), d_yp.get(), d_ybase.get(), d_dr.get(),
                      d_ychk.get(), d_gr.get(), grn, d_itb.get(), rho};
    if (SC.n) {
        P.seg = SC.d_blk.get();
        P.rowseg = SC.d_rowseg.get();
    }
    if (SR.n) {
        D.seg = SR.d_blk.get();
        D.rowseg = SR.d_rowseg.get();
    }
    // the split rows' block pre-passes of iteration offset i of a cycle
    auto pre_primal = [&](int i) {
        if (SC.n)
            launch::split_segments(st, SC.d_items.get(), SC.n, d_cidx_s.get(), d_cval_s.get(), d_y.get(),
                                   d_yp.get(), d_inst.get(), d_itb.get(), i, SC.d_blk.get());
    };
    auto pre_dual = [&](int i) {
        if (SR.n)
            launch::split_segments(st, SR.d_items.get(), SR.n, d_ridx_s.get(), d_rval_s.get(), d_xbar.get(),
                                   nullptr, d_inst.get(), d_itb.get(), i, SR.d_blk.get());
    };
    launch::EnhCheck CK{d_rptr.get(), RD.idx, d_cptr.get(), CD.idx, RD.v64,
                        CD.v64, d_xchk.get(), d_ychk.get(), d_ax.get(), d_lam.get(),
                        d_rl64.get(), d_ru64.get(), d_c64.get(), d_l64.get(), d_u64.get(), d_dc.get(),
                        d_dr.get(), d_gc.get(), gcn, d_gr.get(), grn, d_status.get(), cfg.refine_ratio,
                        cfg.max_refinements, d_xbest.get(), d_lambest.get(),
                        d_ybest.get(), d_axbest.get(), d_xbase.get(), d_ybase.get(), d_cs.get(),
                        d_ls.get(), d_us.get(), d_x.get(), d_x0.get(), d_xp.get(), d_rls.get(),
                        d_rus.get(), d_ybar.get(), d_y.get(), d_y0.get(), d_yp.get(), d_itb.get()};
    const int nsamp = (cfg.profile && !noalloc) ? 256 : 0;
    if (nsamp) {
        while (ctx.events.size() < static_cast<size_t>(4 * nsamp + 4)) {
            hipEvent_t e;
            MPDHG_HIP(hipEventCreate(&e));
            ctx.events.push_back(e);
        }
    }
    // Samples are the first non-check iterations after the first cycle on the
    // full tile lists.  bytes_*_launch is computed for the whole batch, so a
    // sample counts only if every instance was still iterating when it ran:
    // its iteration index is kept and checked against each instance's stop
    // once the loop is over.
    int used = 0;
    std::vector<int> samp_it(static_cast<size_t>(nsamp));
    Pinned pin(2 * static_cast<size_t>(na) * sizeof(int32_t));
    int32_t *slots = static_cast<int32_t *>(pin.p);
    hipEvent_t poll_ev[2] = {nullptr, nullptr};
    if (!noalloc) {
        MPDHG_HIP(hipEventCreateWithFlags(&poll_ev[0], hipEventDisableTiming));
        MPDHG_HIP(hipEventCreateWithFlags(&poll_ev[1], hipEventDisableTiming));
    }
    int it = 0, cycle = 0, covered = na;  // instances (and their nonzeros) the tile lists cover
    int64_t covered_nnz = L.nnz;
    bool finished = false;
    std::vector<DBuf<Tile>> keep_rt, keep_ct;  // compacted tile lists, alive until the end
    std::vector<std::vector<Tile>> host_keep_rt, host_keep_ct;  // sources of their async uploads
    // One check cycle of len iterations in stream order: F fused launch pairs,
    // the FP64 check and decisions, the iteration advance.
    // The fused dual step: one launch, or one per column band (events, when
    // given, at the first band's start and the last band's end).
    auto dual = [&](int i, int check, hipEvent_t e0, hipEvent_t e1) {
        if (bands == 1) {
            launch::enh_dual(st, rts, d_inst.get(), D, i, check, e0, e1);
            return;
        }
        for (int b = 0; b < bands; ++b) {
            const ColumnBands::Band &bd = BD.bd[static_cast<size_t>(b)];
            launch::enh_dual_band(st, BD.sets[static_cast<size_t>(b)], d_inst.get(), D, bd.d_rptr.get(),
                                  bd.dir.idx, bd.dir.v32, BD.hpart.get(), b == 0, b == bands - 1, i, check,
                                  b == 0 ? e0 : nullptr, b == bands - 1 ? e1 : nullptr);
        }
    };
    auto check_step = [&](int len) {
        launch::enh_check(st, rts, cts, d_inst.get(), na, CK, len - 1);
        for (const TileSet &bt : BD.sets) launch::tile_run(st, bt, d_inst.get());
        launch::enh_advance(st, d_itb.get(), len);
    };
    auto issue_cycle = [&](int len) {
        for (int i = 0; i < len; ++i) {
            const int check = (i + 1 == len);
            pre_primal(i);
            launch::enh_primal(st, cts, d_inst.get(), P, i, check);
            pre_dual(i);
            dual(i, check, nullptr, nullptr);
        }
        check_step(len);
    };
    // Full cycles replay a HIP graph recorded from issue_cycle(F); it is
    // recorded again after each tile-list compaction (new lists and grids).
    // A replaced graph may still be executing (the host only waits on the
    // cycle before it), so it is retired, not destroyed: retired graphs are
    // destroyed after the final stream synchronisation.
    struct CycleGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        std::vector<std::pair<hipGraph_t, hipGraphExec_t>> retired;
        void retire() {
            if (e || g) retired.emplace_back(g, e);
            e = nullptr;
            g = nullptr;
        }
        ~CycleGraph() {
            retire();
            for (auto &r : retired) {
                if (r.second) (void)hipGraphExecDestroy(r.second);
                if (r.first) (void)hipGraphDestroy(r.first);
            }
        }
    } graph;
    int lists = 0, graph_lists = -1;  // tile-list version, and the one the graph was recorded for
    try {
        while (it < K && !finished) {
            const int len = std::min(F, K - it);
            if (nsamp && used < nsamp && it >= F && covered == na) {
                // a profiled cycle: eager launches, each fused kernel with events at
                // its own start and end
                for (int i = 0; i < len; ++i) {
                    const int check = (i + 1 == len);
                    const bool samp = used < nsamp && !check;
                    hipEvent_t *e = samp ? ctx.events.data() + 4 * used : nullptr;
                    pre_primal(i);
                    launch::enh_primal(st, cts, d_inst.get(), P, i, check, e ? e[0] : nullptr, e ? e[1] : nullptr);
                    pre_dual(i);
                    dual(i, check, e ? e[2] : nullptr, e ? e[3] : nullptr);
                    if (samp) samp_it[static_cast<size_t>(used++)] = it + i;
                }
                check_step(len);
            } else if (cfg.launch_graph && len == F && !noalloc) {
                if (graph_lists != lists) {
                    graph.retire();
                    MPDHG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                    try {
                        issue_cycle(F);
                    } catch (...) {
                        hipGraph_t partial = nullptr;
                        (void)hipStreamEndCapture(st, &partial);
                        if (partial) (void)hipGraphDestroy(partial);
                        throw;
                    }
                    MPDHG_HIP(hipStreamEndCapture(st, &graph.g));
                    MPDHG_HIP(hipGraphInstantiate(&graph.e, graph.g, nullptr, nullptr, 0));
                    graph_lists = lists;
                }
                MPDHG_HIP(hipGraphLaunch(graph.e, st));
            } else {
                issue_cycle(len);
            }
            it += len;
            MPDHG_HIP(hipGetLastError());
            const int slot = cycle & 1;
            MPDHG_HIP(hipMemcpyAsync(slots + slot * na, d_status.get(), static_cast<size_t>(na) * sizeof(int32_t),
                                     hipMemcpyDeviceToHost, st));
            if (noalloc) MPDHG_HIP(hipStreamSynchronize(st));  // and inspect this cycle at once
            else MPDHG_HIP(hipEventRecord(poll_ev[slot], st));
            if (noalloc || cycle > 0) {  // else inspect the previous cycle while this one runs
                const int prev = noalloc ? slot : slot ^ 1;
                if (!noalloc) MPDHG_HIP(hipEventSynchronize(poll_ev[prev]));
                const int32_t *sv = slots + prev * na;
                const int running = static_cast<int>(std::count(sv, sv + na, kRunning));
                finished = (running == 0);
                int64_t running_nnz = 0;
                for (int a = 0; a < na; ++a)
                    if (sv[a] == kRunning) running_nnz += H[static_cast<size_t>(act[static_cast<size_t>(a)])].nnz;
                if (!finished && 20 * running_nnz <= 19 * covered_nnz) {
                    // Compact once the running instances hold 95 % of the covered
                    // nonzeros or less: later launches cover only their tiles.
                    // Stream order puts the range update after the cycle already
                    // in flight, which still uses the old lists.
                    // the running LPs dealt into the XCD bins again, by size
                    std::vector<int> run;
                    for (int a = 0; a < na; ++a)
                        if (sv[a] == kRunning) run.push_back(a);
                    std::stable_sort(run.begin(), run.end(), [&](int a, int b) {
                        return H[static_cast<size_t>(act[static_cast<size_t>(a)])].nnz >
                               H[static_cast<size_t>(act[static_cast<size_t>(b)])].nnz;
                    });
                    std::vector<Tile> nrt, nct;
                    for (size_t bin = 0; bin < kXcds; ++bin) {
                        for (size_t k = bin; k < run.size(); k += kXcds) {
                            const EnhInst &s0 = hh.inst[static_cast<size_t>(run[k])];
                            nrt.insert(nrt.end(), RD.host.begin() + s0.row_tile_begin, RD.host.begin() + s0.row_tile_end);
                            nct.insert(nct.end(), CD.host.begin() + s0.col_tile_begin, CD.host.begin() + s0.col_tile_end);
                        }
                    }
                    keep_rt.emplace_back();
                    keep_rt.back().upload(nrt, st);
                    keep_ct.emplace_back();
                    keep_ct.back().upload(nct, st);
                    rts = make_tileset(keep_rt.back(), static_cast<int>(nrt.size()), RD.sc, RD.gather);
                    cts = make_tileset(keep_ct.back(), static_cast<int>(nct.size()), CD.sc, CD.gather);
                    rts.stream_nt = cts.stream_nt =
                        stream_bytes([&](int a) { return sv[a] == kRunning; }) > kNtStreamBytes;
                    covered = running;
                    covered_nnz = running_nnz;
                    ++lists;
                    host_keep_rt.push_back(std::move(nrt));
                    host_keep_ct.push_back(std::move(nct));
                }
            }
            ++cycle;
        }
        MPDHG_HIP(hipStreamSynchronize(st));
    } catch (...) {
        for (hipEvent_t e : poll_ev)
            if (e) (void)hipEventDestroy(e);
        throw;
    }
    for (hipEvent_t e : poll_ev)
        if (e) (void)hipEventDestroy(e);
    const double loop_sec = now_sec() - t_loop;

    // ---- results: the best point of each instance ----
    std::vector<double> xb, yb, lb, ab;
    d_xbest.download(xb, C, st);
    d_ybest.download(yb, R, st);
    d_lambest.download(lb, C, st);
    d_axbest.download(ab, R, st);
    MPDHG_HIP(hipMemcpyAsync(hs.data(), d_inst.get(), hs.size() * sizeof(EnhInst), hipMemcpyDeviceToHost, st));
    MPDHG_HIP(hipStreamSynchronize(st));
    // kernel timings of the samples taken while every instance iterated
    // (an instance that stopped after `iterations` iterations ran none past it)
    int first_stop = it;
    for (const EnhInst &s : hs)
        if (s.status != kRunning) first_stop = std::min(first_stop, static_cast<int>(s.iterations));
    double ms_csc = 0, ms_csr = 0;
    int valid = 0;
    for (int s = 0; s < used; ++s) {
        if (samp_it[static_cast<size_t>(s)] >= first_stop) continue;
        float a = 0, b = 0;
        MPDHG_HIP(hipEventElapsedTime(&a, ctx.events[static_cast<size_t>(4 * s)], ctx.events[static_cast<size_t>(4 * s + 1)]));
        MPDHG_HIP(hipEventElapsedTime(&b, ctx.events[static_cast<size_t>(4 * s + 2)], ctx.events[static_cast<size_t>(4 * s + 3)]));
        ms_csc += a;
        ms_csr += b;
        ++valid;
    }
    const double sg = cfg.dual_sign ? -1.0 : 1.0;
    for (int a = 0; a < na; ++a) {
        const int q = act[static_cast<size_t>(a)];
        const HostLp &h = H[static_cast<size_t>(q)];
        mpdhg_result &r = res[q];
        const EnhInst &s = hs[static_cast<size_t>(a)];
        const size_t ro = static_cast<size_t>(L.row_off[static_cast<size_t>(a)]),
                     co = static_cast<size_t>(L.col_off[static_cast<size_t>(a)]);
        r.status = (s.status == kRunning) ? MPDHG_STATUS_ITERATION_LIMIT : s.status;
        r.iterations = (s.status == kRunning) ? it : s.iterations;
        r.restarts = s.restarts;
        r.refinements = s.rounds;
        r.checks = s.checks;
        r.primal_objective = s.best_pobj;
        r.dual_objective = s.best_dobj;
        r.rel_primal = s.best_eps_p;
        r.rel_dual = s.best_eps_d;
        r.rel_gap = s.best_eps_gap;
        r.rel_kkt = s.best_kkt;
        r.step_size = s.eta;
        // the effective weight: a refinement rescales omega by its alpha ratio,
        // which omega alpha_p / alpha_d undoes (a new solve starts at alpha = 1)
        r.primal_weight = s.omega * s.alpha_p / s.alpha_d;
        r.csr_block = bands > 1 ? BD.bd[0].dir.sc.block : RD.sc.block;  // the schedule of A xbar
        r.csr_chunk = bands > 1 ? BD.bd[0].dir.sc.tn : RD.sc.tn;
        r.csr_bands = bands;
        r.csc_block = CD.sc.block;
        r.csc_chunk = CD.sc.tn;
        r.tune_ms = tune_ms;
        if (r.x) std::memcpy(r.x, xb.data() + co, static_cast<size_t>(h.n) * sizeof(double));
        if (r.reduced_costs) std::memcpy(r.reduced_costs, lb.data() + co, static_cast<size_t>(h.n) * sizeof(double));
        if (r.row_activity) std::memcpy(r.row_activity, ab.data() + ro, static_cast<size_t>(h.m) * sizeof(double));
        if (r.y)
            for (int i = 0; i < h.m; ++i) r.y[i] = sg * yb[ro + static_cast<size_t>(i)];
        r.solve_sec = loop_sec;
        r.kernel_ms_spmv_csc = ms_csc;
        r.kernel_ms_spmv_csr = ms_csr;
        r.kernel_launches_csc = r.kernel_launches_csr = valid;
        // K1 streams CSC idx+val, gathers y once, reads c,l,u,x,x0 and writes x,xbar;
        // K2 streams CSR idx+val, gathers xbar once, reads rl,ru,ybar,y,y0, writes y.
        r.bytes_csc_launch = 8.0 * L.nnz + 4.0 * (L.cols + 1) + 4.0 * L.rows + 4.0 * 7 * L.cols;
        r.bytes_csr_launch = 8.0 * L.nnz + 4.0 * (L.rows + 1) + 4.0 * L.cols + 4.0 * 6 * L.rows;
    }
    const double wall = now_sec() - t_call;
    for (int q = 0; q < count; ++q) res[q].time_wall_sec = wall;
}

}  // namespace mpdhg
