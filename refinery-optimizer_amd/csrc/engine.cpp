// Host engine: LP validation and layout, tile planning, and the
// reference-mode driver loop (gpu_solver.cpp:60-307) over a batch.
#include "engine.hpp"

#include <exception>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <thread>

namespace mpdhg {

double now_sec() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

Arena *&thread_arena() {
    thread_local Arena *a = nullptr;
    return a;
}

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw Error(e == hipErrorOutOfMemory ? MPDHG_E_NOMEM : MPDHG_E_HIP,
                    std::string(what) + ": " + hipGetErrorString(e));
}

Context::~Context() {
    for (auto ev : events) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

// Validates the int64 CSC of model::SparseMatrix (sparse_matrix.hpp:38-40)
// and narrows it to int32 -- element by element, which is the copy the
// reference gets wrong at gpu_solver.cpp:202-203 by moving int64 words with
// int32 byte counts.
void narrow_structure(const mpdhg_lp &lp, HostLp &h) {
    const int64_t lim = std::numeric_limits<int32_t>::max() - kTileNnz;
    if (lp.m < 0 || lp.n < 0 || lp.nnz < 0 || lp.m > lim || lp.n > lim || lp.nnz > lim)
        throw Error(MPDHG_E_ARG, "LP dimensions out of range (m, n, nnz must fit int32)");
    h.src = &lp;
    h.m = static_cast<int>(lp.m);
    h.n = static_cast<int>(lp.n);
    h.nnz = static_cast<int>(lp.nnz);
    if (h.n > 0 && (!lp.col_ptr || !lp.c || !lp.col_lower || !lp.col_upper))
        throw Error(MPDHG_E_ARG, "LP column arrays are NULL");
    if (h.m > 0 && (!lp.row_lower || !lp.row_upper))
        throw Error(MPDHG_E_ARG, "LP row-bound arrays are NULL");
    if (h.nnz > 0 && (!lp.row_ind || !lp.vals)) throw Error(MPDHG_E_ARG, "LP nonzero arrays are NULL");
    h.cptr.assign(static_cast<size_t>(h.n) + 1, 0);
    h.cidx.resize(static_cast<size_t>(h.nnz));
    if (h.n > 0) {
        if (lp.col_ptr[0] != 0 || lp.col_ptr[h.n] != lp.nnz)
            throw Error(MPDHG_E_ARG, "col_ptr must start at 0 and end at nnz");
        for (int j = 0; j < h.n; ++j) {
            if (lp.col_ptr[j + 1] < lp.col_ptr[j]) throw Error(MPDHG_E_ARG, "col_ptr not monotone");
            h.cptr[static_cast<size_t>(j) + 1] = static_cast<int>(lp.col_ptr[j + 1]);
        }
    } else if (h.nnz != 0) {
        throw Error(MPDHG_E_ARG, "nonzeros without columns");
    }
    for (int k = 0; k < h.nnz; ++k) {
        const int64_t r = lp.row_ind[k];
        if (r < 0 || r >= lp.m) throw Error(MPDHG_E_ARG, "row index out of range");
        h.cidx[static_cast<size_t>(k)] = static_cast<int>(r);
    }
}

// The CSR by the counting sort of gpu_solver.cpp:92-109 (columns ascending
// within each row).
void build_csr(HostLp &h) {
    h.rptr.assign(static_cast<size_t>(h.m) + 1, 0);
    for (int k = 0; k < h.nnz; ++k) h.rptr[static_cast<size_t>(h.cidx[k]) + 1]++;
    for (int i = 0; i < h.m; ++i) h.rptr[static_cast<size_t>(i) + 1] += h.rptr[static_cast<size_t>(i)];
    std::vector<int> fill(h.rptr.begin(), h.rptr.end());
    h.ridx.resize(static_cast<size_t>(h.nnz));
    h.csc2csr.resize(static_cast<size_t>(h.nnz));
    for (int j = 0; j < h.n; ++j)
        for (int k = h.cptr[j]; k < h.cptr[j + 1]; ++k) {
            const int d = fill[static_cast<size_t>(h.cidx[k])]++;
            h.ridx[static_cast<size_t>(d)] = j;
            h.csc2csr[static_cast<size_t>(k)] = d;
        }
}

// The same CSR on `parts` host threads: the columns are cut into parts
// ranges of about equal nonzeros; part t counts its entries per row, a
// per-row scan over the parts gives part t its first slot in each row, and
// each part scatters its columns in order -- so every row still lists its
// columns ascending, as build_csr does.
void build_csr_parallel(HostLp &h, int parts) {
    const size_t M = static_cast<size_t>(h.m), P = static_cast<size_t>(parts);
    std::vector<int> c0(P + 1);
    for (size_t t = 0; t <= P; ++t) {
        const int64_t want = static_cast<int64_t>(h.nnz) * static_cast<int64_t>(t) / parts;
        c0[t] = static_cast<int>(std::lower_bound(h.cptr.begin(), h.cptr.end(), want) - h.cptr.begin());
    }
    c0[P] = h.n;
    std::vector<int> cnt(P * M, 0);  // entries of part t in row i, then part t's first slot in row i
    parallel_for_each(parts, [&](int t) {
        int *ct = cnt.data() + static_cast<size_t>(t) * M;
        for (int k = h.cptr[static_cast<size_t>(c0[static_cast<size_t>(t)])];
             k < h.cptr[static_cast<size_t>(c0[static_cast<size_t>(t) + 1])]; ++k)
            ct[h.cidx[static_cast<size_t>(k)]]++;
    });
    h.rptr.assign(M + 1, 0);
    for (size_t i = 0; i < M; ++i) {
        int len = 0;
        for (size_t t = 0; t < P; ++t) len += cnt[t * M + i];
        h.rptr[i + 1] = h.rptr[i] + len;
    }
    parallel_for_each(parts, [&](int t) {  // row ranges: first slots
        const size_t i0 = M * static_cast<size_t>(t) / P, i1 = M * (static_cast<size_t>(t) + 1) / P;
        for (size_t i = i0; i < i1; ++i) {
            int at = h.rptr[i];
            for (size_t u = 0; u < P; ++u) {
                const int n = cnt[u * M + i];
                cnt[u * M + i] = at;
                at += n;
            }
        }
    });
    h.ridx.resize(static_cast<size_t>(h.nnz));
    h.csc2csr.resize(static_cast<size_t>(h.nnz));
    parallel_for_each(parts, [&](int t) {
        int *fill = cnt.data() + static_cast<size_t>(t) * M;
        for (int j = c0[static_cast<size_t>(t)]; j < c0[static_cast<size_t>(t) + 1]; ++j)
            for (int k = h.cptr[static_cast<size_t>(j)]; k < h.cptr[static_cast<size_t>(j) + 1]; ++k) {
                const int d = fill[h.cidx[static_cast<size_t>(k)]]++;
                h.ridx[static_cast<size_t>(d)] = j;
                h.csc2csr[static_cast<size_t>(k)] = d;
            }
    });
}

}  // namespace

void prepare_structure(const mpdhg_lp &lp, HostLp &h) {
    narrow_structure(lp, h);
    build_csr(h);
}

// Rows [start, i) form one tile while the tile has at most `block` rows and
// at most tn nonzeros, growing by whole kGroup-row groups; a single denser
// group gets a tile of its own (the kernel streams it in several chunks).
void plan_tiles(const std::vector<int> &ptr, int count, int64_t global_off, int64_t nz_off,
                int64_t grp_off, int inst, int block, int tn, std::vector<Tile> &out) {
    int i = 0;
    while (i < count) {
        const int start = i;
        int nz = 0;
        while (i < count && i - start < block) {
            const int ge = std::min(count, i + kGroup);
            const int gn = ptr[static_cast<size_t>(ge)] - ptr[static_cast<size_t>(i)];
            if (i > start && nz + gn > tn) break;
            nz += gn;
            i = ge;
        }
        out.push_back(Tile{static_cast<int32_t>(global_off + start),
                           static_cast<int32_t>(global_off + i), inst, -1,
                           static_cast<int32_t>(nz_off + ptr[static_cast<size_t>(start)]),
                           static_cast<int32_t>(nz_off + ptr[static_cast<size_t>(i)]), 1,
                           static_cast<int32_t>(grp_off + start / kGroup)});
    }
}

void stratify_by_size(std::vector<int> &act, const std::vector<HostLp> &H) {
    std::vector<int> rank(act);
    std::stable_sort(rank.begin(), rank.end(), [&](int a, int b) {
        return H[static_cast<size_t>(a)].nnz > H[static_cast<size_t>(b)].nnz;
    });
    act.clear();
    for (size_t bin = 0; bin < kXcds; ++bin)
        for (size_t k = bin; k < rank.size(); k += kXcds) act.push_back(rank[k]);
}

void prepare_all(const mpdhg_lp *lps, int count, std::vector<HostLp> &H) {
    H.resize(static_cast<size_t>(count));
    std::vector<std::exception_ptr> err(static_cast<size_t>(count));
    std::vector<int> large;  // LPs whose CSR is built on all host threads, one after the other
    std::mutex mu;
    parallel_for_each(count, [&](int q) {
        try {
            HostLp &h = H[static_cast<size_t>(q)];
            narrow_structure(lps[q], h);
            if (h.nnz >= kParallelCsrNnz) {
                std::lock_guard<std::mutex> lock(mu);
                large.push_back(q);
            } else {
                build_csr(h);
            }
        } catch (...) {
            err[static_cast<size_t>(q)] = std::current_exception();
        }
    });
    for (const std::exception_ptr &e : err)
        if (e) std::rethrow_exception(e);
    const int parts = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    for (int q : large) build_csr_parallel(H[static_cast<size_t>(q)], parts);
}

BatchLayout make_layout(const std::vector<HostLp> &H, const std::vector<int> &act) {
    const size_t na = act.size();
    BatchLayout L;
    L.row_off.resize(na);
    L.col_off.resize(na);
    L.nnz_off.resize(na);
    L.rgrp_off.resize(na);
    L.cgrp_off.resize(na);
    for (size_t a = 0; a < na; ++a) {
        const HostLp &h = H[static_cast<size_t>(act[a])];
        L.row_off[a] = L.rows;
        L.col_off[a] = L.cols;
        L.nnz_off[a] = L.nnz;
        L.rgrp_off[a] = L.rgroups;
        L.cgrp_off[a] = L.cgroups;
        L.rows += h.m;
        L.cols += h.n;
        L.nnz += h.nnz;
        L.rgroups += (h.m + kGroup - 1) / kGroup;
        L.cgroups += (h.n + kGroup - 1) / kGroup;
    }
    const int64_t lim = std::numeric_limits<int32_t>::max() - kTileNnz;
    if (L.rows > lim || L.cols > lim || L.nnz > lim)
        throw Error(MPDHG_E_ARG, "batch too large for int32 indexing; split it");
    return L;
}

TileSet make_tileset(const DBuf<Tile> &d, int count, Schedule sc, bool gather) {
    TileSet ts;
    ts.tiles = d.get();
    ts.count = count;
    ts.chunk = (count + 7) / 8;
    ts.grid = 8 * ts.chunk;
    ts.block = sc.block;
    ts.tn = sc.tn;
    ts.gather = gather;
    return ts;
}

Schedule choose_schedule(int64_t nnz, int64_t rows, int forced_block, int forced_tn) {
    if (forced_block && forced_tn) {
        if (!valid_schedule(forced_block, forced_tn))
            throw Error(MPDHG_E_ARG, "spmv_block / spmv_chunk name no compiled tile schedule");
        return Schedule{forced_block, forced_tn};
    }
    const double avg = static_cast<double>(nnz) / static_cast<double>(std::max<int64_t>(rows, 1));
    Schedule best{0, 0};
    for (int b : {512, 256, 128}) {
        if (forced_block && b != forced_block) continue;
        // the smallest chunk of this workgroup size that holds b rows of
        // average length (the largest one if none does)
        const double need = b * avg;
        Schedule sc{0, 0};
        for (const Schedule &c : kSchedules) {
            if (c.block != b || (forced_tn && c.tn != forced_tn)) continue;
            const bool cf = c.tn >= need, sf = sc.block && sc.tn >= need;
            if (!sc.block || (cf && !sf) || (cf && sf && c.tn < sc.tn) || (!cf && !sf && c.tn > sc.tn))
                sc = c;
        }
        if (!sc.block) continue;
        const int64_t tiles = std::max<int64_t>((rows + b - 1) / b, (nnz + sc.tn - 1) / sc.tn);
        best = sc;
        if (tiles >= kTilesWanted) break;
    }
    if (!best.block) throw Error(MPDHG_E_ARG, "spmv_block / spmv_chunk name no compiled tile schedule");
    return best;
}

int band_count(int forced, int64_t vec_len) {
    if (forced > 0) return std::min<int64_t>(forced, std::max<int64_t>(vec_len, 1));
    const int64_t bytes = 4 * vec_len;
    if (bytes <= kBandMinBytes) return 1;
    return static_cast<int>(std::min<int64_t>(kMaxBands, (bytes + kBandBytes - 1) / kBandBytes));
}

double build_bands(int bands, int rows, int64_t cols, const int *rptr, const int *ridx, const float *rval,
                   const mpdhg_config &cfg, Layout lay, bool tune, float *vin, float *vout,
                   hipStream_t st, ColumnBands &out) {
    const size_t R = static_cast<size_t>(rows);
    std::vector<int64_t> cb(static_cast<size_t>(bands) + 1);
    for (int b = 0; b <= bands; ++b) cb[static_cast<size_t>(b)] = cols * b / bands;
    out.bd = std::vector<ColumnBands::Band>(static_cast<size_t>(bands));
    for (ColumnBands::Band &bd : out.bd) bd.rptr.assign(R + 1, 0);
    // pass 1: nonzeros per row and band; pass 2: scatter in storage order.
    // Both passes are independent per row: ranges of rows run on the host
    // threads.
    const int nr = static_cast<int>(std::min<size_t>(64, std::max<size_t>(1, R / 16384)));
    auto rows_of = [&](int t, size_t &i0, size_t &i1) {
        i0 = R * static_cast<size_t>(t) / static_cast<size_t>(nr);
        i1 = R * (static_cast<size_t>(t) + 1) / static_cast<size_t>(nr);
    };
    parallel_for_each(nr, [&](int t) {
        size_t i0, i1;
        rows_of(t, i0, i1);
        for (size_t i = i0; i < i1; ++i) {
            int b = 0;
            for (int k = rptr[i]; k < rptr[i + 1]; ++k) {
                while (ridx[static_cast<size_t>(k)] >= cb[static_cast<size_t>(b) + 1]) ++b;
                out.bd[static_cast<size_t>(b)].rptr[i + 1]++;
            }
        }
    });
    for (ColumnBands::Band &bd : out.bd) {
        for (size_t i = 0; i < R; ++i) bd.rptr[i + 1] += bd.rptr[i];
        bd.ridx.resize(static_cast<size_t>(bd.rptr[R]));
        bd.rval.resize(static_cast<size_t>(bd.rptr[R]));
    }
    parallel_for_each(nr, [&](int t) {
        size_t i0, i1;
        rows_of(t, i0, i1);
        std::vector<int> fill(static_cast<size_t>(bands));
        for (size_t i = i0; i < i1; ++i) {
            for (int b = 0; b < bands; ++b) fill[static_cast<size_t>(b)] = out.bd[static_cast<size_t>(b)].rptr[i];
            int b = 0;
            for (int k = rptr[i]; k < rptr[i + 1]; ++k) {
                while (ridx[static_cast<size_t>(k)] >= cb[static_cast<size_t>(b) + 1]) ++b;
                ColumnBands::Band &bd = out.bd[static_cast<size_t>(b)];
                const int d = fill[static_cast<size_t>(b)]++;
                bd.ridx[static_cast<size_t>(d)] = ridx[static_cast<size_t>(k)];
                bd.rval[static_cast<size_t>(d)] = rval[static_cast<size_t>(k)];
            }
        }
    });
    double tune_ms = 0.0;
    Schedule bsc{0, 0};
    for (size_t b = 0; b < out.bd.size(); ++b) {
        ColumnBands::Band &bd = out.bd[b];
        bd.d_rptr.upload(bd.rptr, st);
        bd.d_ridx.upload(bd.ridx, st);
        bd.d_rval.upload(bd.rval, st);
        DirPlan P;
        P.ptr = {&bd.rptr};
        P.count = {rows};
        P.row_off = {0};
        P.nnz_off = {0};
        P.grp_off = {0};
        P.vec_off = {cb[b]};
        P.vec_len = {cb[b + 1] - cb[b]};
        P.rows = rows;
        P.nnz = bd.rptr[R];
        const DirSource BS{bd.d_ridx.get(), bd.d_rval.get(), nullptr};
        if (b == 0 && tune) {
            std::vector<TuneTrial> trials;
            tune_ms += tune_dir(P, true, lay, BS, bd.d_rptr.get(), cfg.spmv_block, cfg.spmv_chunk, true,
                                false, vin, vout, nullptr, st, bd.dir, trials);
            bsc = bd.dir.sc;
        } else {
            if (b == 0) bsc = choose_schedule(P.nnz, rows, cfg.spmv_block, cfg.spmv_chunk);
            build_dir(P, bsc, lay, BS, true, false, st, bd.dir);
        }
    }
    out.hpart.resize(R);
    out.sets.clear();
    for (const ColumnBands::Band &bd : out.bd) out.sets.push_back(bd.dir.set());
    return tune_ms;
}

static inline bool finite_bound(double v) { return std::isfinite(v); }

void host_kkt(const mpdhg_lp &lp, const HostLp &h, const std::vector<double> &x,
              const std::vector<double> &y, mpdhg_result &r) {
    const int m = h.m, n = h.n;
    std::vector<double> ax(static_cast<size_t>(m), 0.0);
    for (int j = 0; j < n; ++j)
        for (int k = h.cptr[j]; k < h.cptr[j + 1]; ++k)
            ax[static_cast<size_t>(h.cidx[k])] += lp.vals[k] * x[static_cast<size_t>(j)];
    double bn = 0, cn = 0, rowv = 0, colv = 0, dv = 0, pobj = 0, dobj = 0;
    for (int i = 0; i < m; ++i) {
        const double lo = lp.row_lower[i], hi = lp.row_upper[i], a = ax[static_cast<size_t>(i)];
        if (finite_bound(lo)) bn = std::max(bn, std::fabs(lo));
        if (finite_bound(hi)) bn = std::max(bn, std::fabs(hi));
        double v = 0;
        if (finite_bound(lo) && a < lo) v = lo - a;
        if (finite_bound(hi) && a > hi) v = a - hi;
        rowv = std::max(rowv, v);
        // row-dual sign set: y > 0 needs finite ru, y < 0 finite rl
        const double yr = y[static_cast<size_t>(i)];
        double yi = yr;
        if (!finite_bound(hi) && yi > 0) yi = 0;
        if (!finite_bound(lo) && yi < 0) yi = 0;
        dv = std::max(dv, std::fabs(yr - yi));
        if (yi > 0) dobj -= hi * yi;
        else if (yi < 0) dobj -= lo * yi;
        if (r.row_activity) r.row_activity[i] = a;
    }
    for (int j = 0; j < n; ++j) {
        const double lo = lp.col_lower[j], hi = lp.col_upper[j], xj = x[static_cast<size_t>(j)];
        cn = std::max(cn, std::fabs(lp.c[j]));
        double v = 0;
        if (finite_bound(lo) && xj < lo) v = lo - xj;
        if (finite_bound(hi) && xj > hi) v = xj - hi;
        colv = std::max(colv, v);
        double lam = lp.c[j];
        for (int k = h.cptr[j]; k < h.cptr[j + 1]; ++k)
            lam += lp.vals[k] * y[static_cast<size_t>(h.cidx[k])];
        double p = lam;
        if (!finite_bound(lo) && p > 0) p = 0;
        if (!finite_bound(hi) && p < 0) p = 0;
        dv = std::max(dv, std::fabs(lam - p));
        if (p > 0) dobj += lo * p;
        else if (p < 0) dobj += hi * p;
        pobj += lp.c[j] * xj;
        if (r.reduced_costs) r.reduced_costs[j] = lam;
    }
    r.rel_primal = std::max(rowv, colv) / (1.0 + bn);
    r.rel_dual = dv / (1.0 + cn);
    r.rel_gap = std::fabs(pobj - dobj) / (1.0 + std::fabs(pobj) + std::fabs(dobj));
    r.rel_kkt = (std::isnan(r.rel_primal) || std::isnan(r.rel_dual) || std::isnan(r.rel_gap))
                    ? r.rel_primal + r.rel_dual + r.rel_gap
                    : std::max(r.rel_primal, std::max(r.rel_dual, r.rel_gap));
    r.dual_objective = dobj;
}

namespace {

struct RefHost {
    std::vector<int> cptr, cidx, rptr, ridx;
    std::vector<float> cval, rval, c, l, u, tau, rl, ru, sigma, x0, y0;
    std::vector<RefInst> inst;
};

void event_pool(Context &ctx, size_t n) {
    while (ctx.events.size() < n) {
        hipEvent_t e;
        MPDHG_HIP(hipEventCreate(&e));
        ctx.events.push_back(e);
    }
}

}  // namespace

// Reference-mode batch solve.  Per instance this is gpu_solver.cpp:60-307
// with a correct A^T y; all running instances advance in the same launches.
void solve_reference(Context &ctx, const mpdhg_lp *lps, int count, const mpdhg_config &cfg,
                     const mpdhg_warm_start *warm, mpdhg_result *res) {
    const double t_call = now_sec();
    if (cfg.check_frequency <= 0) throw Error(MPDHG_E_ARG, "check_frequency must be positive");
    std::vector<HostLp> H(static_cast<size_t>(count));
    prepare_all(lps, count, H);
    std::vector<int> act;
    for (int q = 0; q < count; ++q) {
        mpdhg_result &r = res[q];
        r.status = MPDHG_STATUS_UNKNOWN;
        r.iterations = 0;
        r.restarts = r.refinements = r.checks = 0;
        if (H[static_cast<size_t>(q)].m == 0 || H[static_cast<size_t>(q)].n == 0) {
            r.status = MPDHG_STATUS_OPTIMAL;  // gpu_solver.cpp:77-80
            r.primal_objective = 0;
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
    // ---- host preparation (gpu_solver.cpp:82-151, 192-196) ----
    RefHost hh;
    const BatchLayout L = make_layout(H, act);
    const size_t R = static_cast<size_t>(L.rows), C = static_cast<size_t>(L.cols),
                 Z = static_cast<size_t>(L.nnz);
    hh.cptr.resize(C + 1);
    hh.cidx.resize(Z);
    hh.rptr.resize(R + 1);
    hh.ridx.resize(Z);
    hh.cval.resize(Z);
    hh.rval.resize(Z);
    hh.c.resize(C); hh.l.resize(C); hh.u.resize(C); hh.tau.resize(C); hh.x0.resize(C);
    hh.rl.resize(R); hh.ru.resize(R); hh.sigma.resize(R); hh.y0.assign(R, 0.0f);
    hh.inst.resize(static_cast<size_t>(na));
    for (int a = 0; a < na; ++a) {
        const int q = act[static_cast<size_t>(a)];
        const HostLp &h = H[static_cast<size_t>(q)];
        const mpdhg_lp &lp = lps[q];
        const int64_t ro = L.row_off[static_cast<size_t>(a)], co = L.col_off[static_cast<size_t>(a)],
                      zo = L.nnz_off[static_cast<size_t>(a)];
        for (int j = 0; j <= h.n; ++j) hh.cptr[static_cast<size_t>(co + j)] = static_cast<int>(zo + h.cptr[j]);
        for (int i = 0; i <= h.m; ++i) hh.rptr[static_cast<size_t>(ro + i)] = static_cast<int>(zo + h.rptr[i]);
        for (int k = 0; k < h.nnz; ++k) {
            const float v = static_cast<float>(lp.vals[k]);  // :88-89
            hh.cval[static_cast<size_t>(zo + k)] = v;
            hh.cidx[static_cast<size_t>(zo + k)] = static_cast<int>(ro + h.cidx[k]);
            const int d = h.csc2csr[static_cast<size_t>(k)];
            hh.rval[static_cast<size_t>(zo + d)] = v;
            hh.ridx[static_cast<size_t>(zo + d)] = static_cast<int>(co + h.ridx[d]);
        }
        float mac = 1.0f;  // :113-119
        for (int j = 0; j < h.n; ++j) {
            const size_t g = static_cast<size_t>(co + j);
            hh.c[g] = static_cast<float>(lp.c[j]);
            hh.l[g] = static_cast<float>(lp.col_lower[j]);
            hh.u[g] = static_cast<float>(lp.col_upper[j]);
            mac = std::max(mac, std::abs(hh.c[g]));
        }
        float mab = 1.0f;  // :120-126
        for (int i = 0; i < h.m; ++i) {
            const size_t g = static_cast<size_t>(ro + i);
            hh.rl[g] = static_cast<float>(lp.row_lower[i]);
            hh.ru[g] = static_cast<float>(lp.row_upper[i]);
            if (hh.ru[g] < 1e15f) mab = std::max(mab, std::abs(hh.ru[g]));
            if (hh.rl[g] > -1e15f) mab = std::max(mab, std::abs(hh.rl[g]));
        }
        const double *wx = (warm && warm[q].x) ? warm[q].x : nullptr;
        for (int j = 0; j < h.n; ++j) {  // :130-140
            float s = 0.0f;
            for (int k = h.cptr[j]; k < h.cptr[j + 1]; ++k) s += std::abs(hh.cval[static_cast<size_t>(zo + k)]);
            const size_t g = static_cast<size_t>(co + j);
            hh.tau[g] = (cfg.step_tau > 0.0f) ? cfg.step_tau : 0.95f / std::max(1.0f, s);
            // :194-196, or the warm start clamped to the bounds
            const float x0 = wx ? static_cast<float>(wx[j]) : 0.0f;
            hh.x0[g] = std::max(hh.l[g], std::min(hh.u[g], x0));
        }
        if (warm && warm[q].y) {
            const double sg = cfg.dual_sign ? -1.0 : 1.0;
            for (int i = 0; i < h.m; ++i) hh.y0[static_cast<size_t>(ro + i)] = static_cast<float>(sg * warm[q].y[i]);
        }
        for (int i = 0; i < h.m; ++i) {  // :141-151
            float s = 0.0f;
            for (int k = h.rptr[i]; k < h.rptr[i + 1]; ++k) s += std::abs(hh.rval[static_cast<size_t>(zo + k)]);
            hh.sigma[static_cast<size_t>(ro + i)] =
                (cfg.step_sigma > 0.0f) ? cfg.step_sigma : 0.95f / std::max(1.0f, s);
        }
        RefInst &s = hh.inst[static_cast<size_t>(a)];
        std::memset(&s, 0, sizeof(s));
        s.status = kRunning;
        s.best_res = 1e30f;
        s.max_abs_c = mac;
        s.max_abs_b = mab;
    }
    // ---- device buffers and upload (gpu_solver.cpp:153-216) ----
    hipStream_t st = ctx.stream;
    DBuf<int> d_cptr, d_cidx_s, d_rptr, d_ridx_s;
    DBuf<float> d_cval_s, d_rval_s, d_c, d_l, d_u, d_tau, d_rl, d_ru, d_sigma;
    DBuf<float> d_x(C), d_xbar(C), d_xbest(C), d_y(R);
    DBuf<RefInst> d_inst;
    // Under an arena (the no-allocation entry point) the solve allocates
    // nothing beyond it: no schedule search, no profiling events, no graph,
    // and status polls that wait on the stream instead of on events.
    const bool noalloc = arena_active();
    d_cptr.upload(hh.cptr, st); d_cidx_s.upload(hh.cidx, st); d_cval_s.upload(hh.cval, st);
    d_rptr.upload(hh.rptr, st); d_ridx_s.upload(hh.ridx, st); d_rval_s.upload(hh.rval, st);
    d_c.upload(hh.c, st); d_l.upload(hh.l, st); d_u.upload(hh.u, st); d_tau.upload(hh.tau, st);
    d_rl.upload(hh.rl, st); d_ru.upload(hh.ru, st); d_sigma.upload(hh.sigma, st);
    // ---- tile schedules: searched on this batch's own matrix (the iterate
    // buffers serve as scratch before they are initialised), or by rule ----
    const Layout lay = static_cast<Layout>(cfg.spmv_layout);
    const DirPlan RP = dir_plan(H, act, L, true), CP = dir_plan(H, act, L, false);
    const DirSource RS{d_ridx_s.get(), d_rval_s.get(), nullptr}, CS{d_cidx_s.get(), d_cval_s.get(), nullptr};
    DevDir RD, CD;
    double tune_ms = 0.0;
    if (tune_wanted(cfg, L.nnz) && !noalloc) {
        std::vector<TuneTrial> trials;
        DBuf<double> d_gtmp(static_cast<size_t>(L.cgroups));
        tune_ms += tune_dir(RP, true, lay, RS, d_rptr.get(), cfg.spmv_block, cfg.spmv_chunk, true, false,
                            d_x.get(), d_y.get(), nullptr, st, RD, trials);
        tune_ms += tune_dir(CP, false, lay, CS, d_cptr.get(), cfg.spmv_block, cfg.spmv_chunk, true, false,
                            d_y.get(), d_x.get(), d_gtmp.get(), st, CD, trials);
    } else {
        build_dir(RP, choose_schedule(L.nnz, L.rows, cfg.spmv_block, cfg.spmv_chunk), lay, RS, true, false, st, RD);
        build_dir(CP, choose_schedule(L.nnz, L.cols, cfg.spmv_block, cfg.spmv_chunk), lay, CS, true, false, st, CD);
    }
    // ---- column bands of the A xbar product (one LP whose xbar exceeds an
    // XCD's L2; the no-allocation solve keeps the unbanded product, whose
    // results are the same bit for bit) ----
    const int bands = (na == 1 && !noalloc) ? band_count(cfg.spmv_bands, L.cols) : 1;
    ColumnBands BD;
    if (bands > 1)
        tune_ms += build_bands(bands, static_cast<int>(R), L.cols, hh.rptr.data(), hh.ridx.data(), hh.rval.data(), cfg, lay,
                               tune_wanted(cfg, L.nnz) && !noalloc, d_x.get(), d_y.get(), st, BD);
    for (int a = 0; a < na; ++a) {
        RefInst &s = hh.inst[static_cast<size_t>(a)];
        s.row_tile_begin = RD.first[static_cast<size_t>(a)];
        s.row_tile_end = RD.first[static_cast<size_t>(a) + 1];
        s.col_tile_begin = CD.first[static_cast<size_t>(a)];
        s.col_tile_end = CD.first[static_cast<size_t>(a) + 1];
    }
    d_inst.upload(hh.inst, st);
    MPDHG_HIP(hipMemcpyAsync(d_x.get(), hh.x0.data(), C * sizeof(float), hipMemcpyHostToDevice, st));
    MPDHG_HIP(hipMemcpyAsync(d_xbar.get(), hh.x0.data(), C * sizeof(float), hipMemcpyHostToDevice, st));
    MPDHG_HIP(hipMemcpyAsync(d_xbest.get(), hh.x0.data(), C * sizeof(float), hipMemcpyHostToDevice, st));
    MPDHG_HIP(hipMemcpyAsync(d_y.get(), hh.y0.data(), R * sizeof(float), hipMemcpyHostToDevice, st));
    DBuf<float> d_prow(RD.host.size()), d_pcol(CD.host.size());
    const TileSet rts = RD.set(), cts = CD.set();
    MPDHG_HIP(hipStreamSynchronize(st));

    // ---- profiling samples: HIP events around the fused SpMV launches ----
    const int K = cfg.max_iterations, F = cfg.check_frequency;
    const int nsamp = (cfg.profile && !noalloc) ? std::min(256, std::max(K, 0)) : 0;
    if (nsamp) event_pool(ctx, static_cast<size_t>(4 * nsamp));
    int used = 0;

    // ---- hot loop (gpu_solver.cpp:227-271) ----
    // The host reads each check's status snapshot one check late, while the
    // next F iterations are already queued: the kernels skip instances that
    // have stopped, so iterations queued after the last instance stopped
    // change nothing, and the device never drains between checks.
    std::vector<RefInst> hs(static_cast<size_t>(na));
    Pinned snap(2 * hs.size() * sizeof(RefInst));
    RefInst *const snaps[2] = {static_cast<RefInst *>(snap.p), static_cast<RefInst *>(snap.p) + na};
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() { for (auto v : e) if (v) (void)hipEventDestroy(v); }
    } ev;
    if (!noalloc)
        for (auto &v : ev.e) MPDHG_HIP(hipEventCreateWithFlags(&v, hipEventDisableTiming));
    // The iteration count lives on the device (advanced by the last launch of
    // each check cycle), so one recorded cycle -- F fused pairs, the check,
    // the decisions -- replays unchanged as a HIP graph for every full cycle.
    DBuf<int> d_itb(1);
    MPDHG_HIP(hipMemsetAsync(d_itb.get(), 0, sizeof(int), st));
    auto issue_cycle = [&](int len, bool sample) {
        for (int i = 0; i < len; ++i) {
            // sampled pairs: events at each fused kernel's own start and end
            const bool samp = sample && used < nsamp;
            hipEvent_t *e = samp ? ctx.events.data() + 4 * used : nullptr;
            launch::ref_primal(st, cts, d_inst.get(), d_cptr.get(), CD.idx, CD.v32, d_y.get(), d_c.get(),
                               d_l.get(), d_u.get(), d_tau.get(), d_x.get(), d_xbar.get(), e ? e[0] : nullptr,
                               e ? e[1] : nullptr);
            if (bands > 1) {
                for (int b = 0; b < bands; ++b) {
                    const ColumnBands::Band &bd = BD.bd[static_cast<size_t>(b)];
                    launch::ref_dual_band(st, BD.sets[static_cast<size_t>(b)], d_inst.get(), bd.d_rptr.get(),
                                          bd.dir.idx, bd.dir.v32, d_xbar.get(), BD.hpart.get(), b == 0,
                                          b == bands - 1, d_rl.get(), d_ru.get(), d_sigma.get(), d_y.get(),
                                          (e && b == 0) ? e[2] : nullptr,
                                          (e && b == bands - 1) ? e[3] : nullptr);
                }
            } else {
                launch::ref_dual(st, rts, d_inst.get(), d_rptr.get(), RD.idx, RD.v32, d_xbar.get(), d_rl.get(),
                                 d_ru.get(), d_sigma.get(), d_y.get(), e ? e[2] : nullptr, e ? e[3] : nullptr);
            }
            if (samp) ++used;
            // (:237-239 restart kernel omitted: it writes xbar = x, and the next
            //  primal update overwrites xbar before anything reads it.)
        }
        launch::ref_rowres(st, rts, d_inst.get(), d_rptr.get(), RD.idx, RD.v32, d_x.get(), d_rl.get(),
                           d_ru.get(), d_prow.get());
        launch::ref_colres(st, cts, d_inst.get(), d_cptr.get(), CD.idx, CD.v32, d_y.get(), d_x.get(),
                           d_c.get(), d_l.get(), d_u.get(), d_tau.get(), d_pcol.get());
        launch::ref_finalize(st, d_inst.get(), na, d_prow.get(), d_pcol.get(), d_itb.get(), len, K,
                             cfg.tol_primal, cfg.tol_dual);
        launch::ref_copy_best(st, cts, d_inst.get(), d_x.get(), d_xbest.get());
        launch::tile_run(st, rts, d_inst.get());
        launch::tile_run(st, cts, d_inst.get());
        for (const TileSet &bt : BD.sets) launch::tile_run(st, bt, d_inst.get());
        launch::enh_advance(st, d_itb.get(), len);
    };
    struct CycleGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        ~CycleGraph() {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
        }
    } graph;
    int nchk = 0;
    const double t_loop = now_sec();
    for (int done = 0; done < K;) {
        const int len = std::min(F, K - done);
        if (nsamp && used < nsamp) {
            issue_cycle(len, true);  // profiled cycles: eager, with events around the fused pairs
        } else if (cfg.launch_graph && len == F && !noalloc) {
            if (!graph.e) {
                MPDHG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                try {
                    issue_cycle(F, false);
                } catch (...) {
                    hipGraph_t partial = nullptr;
                    (void)hipStreamEndCapture(st, &partial);
                    if (partial) (void)hipGraphDestroy(partial);
                    throw;
                }
                MPDHG_HIP(hipStreamEndCapture(st, &graph.g));
                MPDHG_HIP(hipGraphInstantiate(&graph.e, graph.g, nullptr, nullptr, 0));
            }
            MPDHG_HIP(hipGraphLaunch(graph.e, st));
        } else {
            issue_cycle(len, false);
        }
        done += len;
        MPDHG_HIP(hipGetLastError());
        const int cur = nchk & 1;
        MPDHG_HIP(hipMemcpyAsync(snaps[cur], d_inst.get(), hs.size() * sizeof(RefInst), hipMemcpyDeviceToHost, st));
        if (noalloc) {  // this check's status, at once
            MPDHG_HIP(hipStreamSynchronize(st));
            ++nchk;
            bool any = false;
            for (int a = 0; a < na; ++a) any = any || (snaps[cur][a].status == kRunning);
            if (!any) break;
            continue;
        }
        MPDHG_HIP(hipEventRecord(ev.e[cur], st));
        if (nchk++ > 0) {
            MPDHG_HIP(hipEventSynchronize(ev.e[cur ^ 1]));
            bool any = false;
            for (int a = 0; a < na; ++a) any = any || (snaps[cur ^ 1][a].status == kRunning);
            if (!any) break;
        }
    }
    MPDHG_HIP(hipStreamSynchronize(st));
    const double loop_sec = now_sec() - t_loop;
    double ms_csc = 0, ms_csr = 0;
    for (int s = 0; s < used; ++s) {
        float a = 0, b = 0;
        MPDHG_HIP(hipEventElapsedTime(&a, ctx.events[static_cast<size_t>(4 * s)], ctx.events[static_cast<size_t>(4 * s + 1)]));
        MPDHG_HIP(hipEventElapsedTime(&b, ctx.events[static_cast<size_t>(4 * s + 2)], ctx.events[static_cast<size_t>(4 * s + 3)]));
        ms_csc += a;
        ms_csr += b;
    }
    MPDHG_HIP(hipMemcpyAsync(hs.data(), d_inst.get(), hs.size() * sizeof(RefInst),
                             hipMemcpyDeviceToHost, st));
    std::vector<float> xb, yv;
    d_xbest.download(xb, C, st);
    d_y.download(yv, R, st);
    MPDHG_HIP(hipStreamSynchronize(st));

    // ---- results (gpu_solver.cpp:293-305) ----
    for (int a = 0; a < na; ++a) {
        const int q = act[static_cast<size_t>(a)];
        const HostLp &h = H[static_cast<size_t>(q)];
        const mpdhg_lp &lp = lps[q];
        mpdhg_result &r = res[q];
        const RefInst &s = hs[static_cast<size_t>(a)];
        const int64_t ro = L.row_off[static_cast<size_t>(a)], co = L.col_off[static_cast<size_t>(a)];
        if (s.status == kRunning) {  // K <= 0: the reference loop never runs
            r.status = MPDHG_STATUS_ITERATION_LIMIT;
            r.iterations = 1;
        } else {
            r.status = s.status;
            r.iterations = s.iterations;
        }
        r.checks = s.checks;
        r.ref_primal = s.last_p;
        r.ref_dual = s.last_d;
        std::vector<double> x(static_cast<size_t>(h.n)), y(static_cast<size_t>(h.m));
        double obj = 0.0;
        for (int j = 0; j < h.n; ++j) {
            x[static_cast<size_t>(j)] = static_cast<double>(xb[static_cast<size_t>(co + j)]);
            obj += lp.c[j] * x[static_cast<size_t>(j)];
        }
        for (int i = 0; i < h.m; ++i) y[static_cast<size_t>(i)] = static_cast<double>(yv[static_cast<size_t>(ro + i)]);
        r.primal_objective = obj;
        if (r.x) std::memcpy(r.x, x.data(), x.size() * sizeof(double));
        if (r.y) {
            const double sg = cfg.dual_sign ? -1.0 : 1.0;
            for (int i = 0; i < h.m; ++i) r.y[i] = sg * y[static_cast<size_t>(i)];
        }
        host_kkt(lp, h, x, y, r);
        r.solve_sec = loop_sec;
        r.kernel_ms_spmv_csc = ms_csc;
        r.kernel_ms_spmv_csr = ms_csr;
        r.kernel_launches_csc = r.kernel_launches_csr = used;
        r.bytes_csc_launch = 8.0 * L.nnz + 4.0 * (L.cols + 1) + 4.0 * L.rows + 4.0 * 7 * L.cols;
        r.bytes_csr_launch = 8.0 * L.nnz + 4.0 * (L.rows + 1) + 4.0 * L.cols + 4.0 * 5 * L.rows;
    }
    for (int a = 0; a < na; ++a) {
        mpdhg_result &r = res[act[static_cast<size_t>(a)]];
        r.csr_block = bands > 1 ? BD.bd[0].dir.sc.block : RD.sc.block;  // the schedule of A xbar
        r.csr_chunk = bands > 1 ? BD.bd[0].dir.sc.tn : RD.sc.tn;
        r.csc_block = CD.sc.block;
        r.csc_chunk = CD.sc.tn;
        r.tune_ms = tune_ms;
        r.csr_bands = bands;
    }
    const double wall = now_sec() - t_call;
    for (int q = 0; q < count; ++q) res[q].time_wall_sec = wall;
}

}  // namespace mpdhg
