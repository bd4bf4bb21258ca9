// Host side of the tile layouts: the per-direction tile plan, the device-
// built gather order, and the schedule search that times every compiled
// tile schedule on the batch's own matrix.
#include "engine.hpp"

#include <algorithm>
#include <limits>

namespace mpdhg {

DirPlan dir_plan(const std::vector<HostLp> &H, const std::vector<int> &act, const BatchLayout &L,
                 bool csr) {
    const size_t na = act.size();
    DirPlan P;
    P.ptr.resize(na);
    P.count.resize(na);
    P.vec_len.resize(na);
    for (size_t a = 0; a < na; ++a) {
        const HostLp &h = H[static_cast<size_t>(act[a])];
        P.ptr[a] = csr ? &h.rptr : &h.cptr;
        P.count[a] = csr ? h.m : h.n;
        P.vec_len[a] = csr ? h.n : h.m;
    }
    P.row_off = csr ? L.row_off : L.col_off;
    P.grp_off = csr ? L.rgrp_off : L.cgrp_off;
    P.vec_off = csr ? L.col_off : L.row_off;
    P.nnz_off = L.nnz_off;
    P.rows = csr ? L.rows : L.cols;
    P.nnz = L.nnz;
    return P;
}

void build_dir(const DirPlan &P, Schedule sc, Layout layout, const DirSource &src, bool want32,
               bool want64, hipStream_t st, DevDir &out) {
    bool gather = (layout == Layout::Auto);
    for (int64_t len : P.vec_len) gather = gather && (len < gather_local_limit(sc.tn));
    out.sc = sc;
    out.gather = gather;
    out.host.clear();
    out.first.assign(P.ptr.size() + 1, 0);
    for (size_t a = 0; a < P.ptr.size(); ++a) {
        const size_t first = out.host.size();
        out.first[a] = static_cast<int32_t>(first);
        plan_tiles(*P.ptr[a], P.count[a], P.row_off[a], P.nnz_off[a], P.grp_off[a], static_cast<int>(a),
                   sc.block, sc.tn, out.host);
        if (gather)
            for (size_t t = first; t < out.host.size(); ++t) out.host[t].base = static_cast<int32_t>(P.vec_off[a]);
    }
    out.first[P.ptr.size()] = static_cast<int32_t>(out.host.size());
    out.tiles.upload(out.host, st);
    if (!gather) {
        out.idx = src.idx;
        out.v32 = want32 ? src.v32 : nullptr;
        out.v64 = want64 ? src.v64 : nullptr;
        return;
    }
    const size_t z = static_cast<size_t>(std::max<int64_t>(P.nnz, 1));
    if (out.own_idx.size() < z) out.own_idx.resize(z);
    if (want32 && out.own32.size() < z) out.own32.resize(z);
    if (want64 && out.own64.size() < z) out.own64.resize(z);
    out.idx = out.own_idx.get();
    out.v32 = want32 ? out.own32.get() : nullptr;
    out.v64 = want64 ? out.own64.get() : nullptr;
    launch::gather_order(st, out.set(), src.idx, out.own_idx.get(), want32 ? src.v32 : nullptr,
                         want32 ? out.own32.get() : nullptr, want64 ? src.v64 : nullptr,
                         want64 ? out.own64.get() : nullptr);
    MPDHG_HIP(hipGetLastError());
}

bool tune_wanted(const mpdhg_config &cfg, int64_t nnz) {
    if (cfg.spmv_block && cfg.spmv_chunk) return false;
    return cfg.spmv_tune == 2 || (cfg.spmv_tune == 1 && nnz >= kTuneMinNnz);
}

double tune_dir(const DirPlan &P, bool csr, Layout layout, const DirSource &src, const int *ptr,
                int forced_block, int forced_tn, bool want32, bool want64, float *vin, float *vout,
                double *gpart, hipStream_t st, DevDir &out, std::vector<TuneTrial> &trials) {
    const double t0 = now_sec();
    constexpr int kReps = 3;
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() { for (auto v : e) if (v) (void)hipEventDestroy(v); }
    } ev;
    for (auto &v : ev.e) MPDHG_HIP(hipEventCreate(&v));
    DevDir cand[2];  // the best so far and the one being tried
    int best = -1;
    double best_us = std::numeric_limits<double>::infinity();
    trials.clear();
    for (const Schedule &sc : kSchedules) {
        if ((forced_block && sc.block != forced_block) || (forced_tn && sc.tn != forced_tn)) continue;
        const int slot = best == 0 ? 1 : 0;
        DevDir &d = cand[slot];
        build_dir(P, sc, layout, src, true, false, st, d);
        const TileSet ts = d.set();
        auto spmv = [&] {
            if (csr)
                launch::pow_rows(st, ts, ptr, d.idx, d.v32, vin, vout);
            else
                launch::pow_cols(st, ts, ptr, d.idx, d.v32, vin, vout, gpart);
        };
        spmv();  // warm-up (instruction cache, TLB)
        MPDHG_HIP(hipEventRecord(ev.e[0], st));
        for (int r = 0; r < kReps; ++r) spmv();
        MPDHG_HIP(hipEventRecord(ev.e[1], st));
        MPDHG_HIP(hipEventSynchronize(ev.e[1]));
        float ms = 0.0f;
        MPDHG_HIP(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
        const double us = 1e3 * ms / kReps;
        trials.push_back(TuneTrial{sc, d.gather, us});
        if (us < best_us) {
            best_us = us;
            best = slot;
        }
    }
    if (best < 0) throw Error(MPDHG_E_ARG, "spmv_block / spmv_chunk name no compiled tile schedule");
    // the winner, with the value copies the solve reads
    build_dir(P, cand[best].sc, layout, src, want32, want64, st, out);
    MPDHG_HIP(hipStreamSynchronize(st));
    return 1e3 * (now_sec() - t0);
}

}  // namespace mpdhg
