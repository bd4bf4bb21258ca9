// Synthetic LP workloads: seeded, portable generators (xoshiro256** seeded
// through SplitMix64) so that the same instances can be regenerated from
// Python for small sizes.  The shapes follow the workloads the engine is
// measured on: random equality LPs (m rows, n columns, k nonzeros per
// column) and refinery staircase LPs (T time blocks).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "mpdhg.h"
#include "mpdhg_internal.hpp"

namespace mpdhg {

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &st) {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t st = seed;
        for (auto &w : s) w = splitmix(st);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double u01() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    double uni(double a, double b) { return a + (b - a) * u01(); }
    uint64_t below(uint64_t k) { return next() % k; }
    double normal() {  // Box-Muller, cosine branch only
        double u1 = u01();
        const double u2 = u01();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    // Floyd's sampling of k distinct values of [0, n), returned ascending.
    std::vector<int64_t> sample(int64_t n, int64_t k) {
        std::vector<int64_t> out;
        out.reserve(static_cast<size_t>(k));
        for (int64_t t = n - k; t < n; ++t) {
            const int64_t r = static_cast<int64_t>(below(static_cast<uint64_t>(t + 1)));
            if (std::find(out.begin(), out.end(), r) != out.end())
                out.push_back(t);
            else
                out.push_back(r);
        }
        std::sort(out.begin(), out.end());
        return out;
    }
};

// Floyd sample for large k (set membership through a bitmap).
std::vector<int64_t> sample_big(Rng &g, int64_t n, int64_t k) {
    std::vector<uint8_t> in(static_cast<size_t>(n), 0);
    std::vector<int64_t> out;
    out.reserve(static_cast<size_t>(k));
    for (int64_t t = n - k; t < n; ++t) {
        const int64_t r = static_cast<int64_t>(g.below(static_cast<uint64_t>(t + 1)));
        const int64_t pick = in[static_cast<size_t>(r)] ? t : r;
        in[static_cast<size_t>(pick)] = 1;
        out.push_back(pick);
    }
    std::sort(out.begin(), out.end());
    return out;
}

}  // namespace

void finish_views(OwnedLp *o) {
    o->view.m = o->m;
    o->view.n = o->n;
    o->view.nnz = static_cast<int64_t>(o->vals.size());
    o->view.col_ptr = o->col_ptr.data();
    o->view.row_ind = o->row_ind.data();
    o->view.vals = o->vals.data();
    o->view.c = o->c.data();
    o->view.col_lower = o->l.data();
    o->view.col_upper = o->u.data();
    o->view.row_lower = o->rl.data();
    o->view.row_upper = o->ru.data();
    o->row_name_ptr.clear();
    for (const auto &nm : o->row_names) o->row_name_ptr.push_back(nm.c_str());
    o->col_name_ptr.clear();
    for (const auto &nm : o->col_names) o->col_name_ptr.push_back(nm.c_str());
    mpdhg_lp_meta &e = o->meta;
    e.name = o->name.c_str();
    e.sense = o->sense;
    e.obj_offset = o->obj_offset;
    e.row_names = o->row_names.empty() ? nullptr : o->row_name_ptr.data();
    e.col_names = o->col_names.empty() ? nullptr : o->col_name_ptr.data();
    e.var_type = o->var_type.empty() ? nullptr : o->var_type.data();
    e.has_quadratic = o->has_q ? 1 : 0;
    e.q_nnz = static_cast<int64_t>(o->q_vals.size());
    e.q_col_ptr = o->has_q ? o->q_col_ptr.data() : nullptr;
    e.q_row_ind = o->has_q ? o->q_row_ind.data() : nullptr;
    e.q_vals = o->has_q ? o->q_vals.data() : nullptr;
}

OwnedLp *gen_random(int64_t m, int64_t n, int32_t k, uint64_t seed) {
    auto *o = new OwnedLp();
    o->m = m;
    o->n = n;
    Rng g(seed);
    const int64_t kk = std::min<int64_t>(k, m);
    o->col_ptr.assign(static_cast<size_t>(n + 1), 0);
    o->row_ind.reserve(static_cast<size_t>(n * kk));
    o->vals.reserve(static_cast<size_t>(n * kk));
    for (int64_t j = 0; j < n; ++j) {
        std::vector<int64_t> rows = g.sample(m, kk);
        for (int64_t r : rows) {
            o->row_ind.push_back(r);
            o->vals.push_back(g.uni(-1.0, 1.0));
        }
        o->col_ptr[static_cast<size_t>(j + 1)] = static_cast<int64_t>(o->row_ind.size());
    }
    std::vector<double> x0(static_cast<size_t>(n));
    for (auto &v : x0) v = g.uni(0.0, 10.0);
    std::vector<double> b(static_cast<size_t>(m), 0.0);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = o->col_ptr[j]; p < o->col_ptr[j + 1]; ++p)
            b[static_cast<size_t>(o->row_ind[p])] += o->vals[p] * x0[static_cast<size_t>(j)];
    std::vector<double> y0(static_cast<size_t>(m));
    for (auto &v : y0) v = g.normal();
    std::vector<double> s(static_cast<size_t>(n));
    for (auto &v : s) v = g.u01();
    for (int64_t j : sample_big(g, n, (n * 4) / 10)) s[static_cast<size_t>(j)] = 0.0;
    o->c.resize(static_cast<size_t>(n));
    for (int64_t j = 0; j < n; ++j) {
        double acc = 0.0;
        for (int64_t p = o->col_ptr[j]; p < o->col_ptr[j + 1]; ++p)
            acc += o->vals[p] * y0[static_cast<size_t>(o->row_ind[p])];
        o->c[static_cast<size_t>(j)] = acc + s[static_cast<size_t>(j)];
    }
    o->l.assign(static_cast<size_t>(n), 0.0);
    o->u.assign(static_cast<size_t>(n), 10.0);
    o->rl = b;
    o->ru = b;
    finish_views(o);
    return o;
}

OwnedLp *gen_staircase(int32_t T, int32_t R, int32_t C, uint64_t seed) {
    const double inf = std::numeric_limits<double>::infinity();
    const int32_t P = R / 2, Q = (3 * R) / 10, U = R - P - Q, NC = C - 2 * P;
    if (T < 1 || R < 10 || NC < 1) return nullptr;
    auto *o = new OwnedLp();
    o->m = static_cast<int64_t>(T) * R;
    o->n = static_cast<int64_t>(T) * C;
    o->row_kind.assign(static_cast<size_t>(o->m), 0);
    o->col_kind.assign(static_cast<size_t>(o->n), 0);
    o->row_block.resize(static_cast<size_t>(o->m));
    o->col_block.resize(static_cast<size_t>(o->n));
    for (int64_t i = 0; i < o->m; ++i) o->row_block[static_cast<size_t>(i)] = static_cast<int32_t>(i / R);
    for (int64_t j = 0; j < o->n; ++j) {
        const int64_t k = j % C;
        o->col_block[static_cast<size_t>(j)] = static_cast<int32_t>(j / C);
        o->col_kind[static_cast<size_t>(j)] = static_cast<int8_t>(k < NC ? 0 : (k < NC + P ? 1 : 2));
    }
    o->rl.assign(static_cast<size_t>(o->m), 0.0);
    o->ru.assign(static_cast<size_t>(o->m), 0.0);
    o->c.assign(static_cast<size_t>(o->n), 0.0);
    o->l.assign(static_cast<size_t>(o->n), 0.0);
    o->u.assign(static_cast<size_t>(o->n), 0.0);
    o->col_ptr.assign(static_cast<size_t>(o->n + 1), 0);
    Rng g(seed);
    const int32_t blend_k = std::min(10, P);
    for (int32_t t = 0; t < T; ++t) {
        const int64_t r0 = static_cast<int64_t>(t) * R, c0 = static_cast<int64_t>(t) * C;
        // blend-quality rows: members and coefficients, row-wise
        std::vector<std::vector<std::pair<int32_t, double>>> sale_blend(static_cast<size_t>(P));
        for (int32_t q = 0; q < Q; ++q) {
            std::vector<int64_t> mem = g.sample(P, blend_k);
            for (int64_t k : mem) sale_blend[static_cast<size_t>(k)].push_back({q, g.uni(0.5, 1.5)});
            const int64_t row = r0 + P + q;
            o->row_kind[static_cast<size_t>(row)] = 1;
            o->rl[static_cast<size_t>(row)] = -inf;
            o->ru[static_cast<size_t>(row)] = g.uni(200.0, 2000.0);
        }
        for (int32_t q = 0; q < U; ++q) {
            const int64_t row = r0 + P + Q + q;
            o->row_kind[static_cast<size_t>(row)] = 2;
            o->rl[static_cast<size_t>(row)] = -inf;
            o->ru[static_cast<size_t>(row)] = g.uni(500.0, 3000.0);
        }
        // crude columns: yields into 4..8 product balances (sum <= 1), one unit
        for (int32_t cc = 0; cc < NC; ++cc) {
            const int64_t col = c0 + cc;
            const int64_t nk = std::min<int64_t>(4 + static_cast<int64_t>(g.below(5)), P);
            std::vector<int64_t> prods = g.sample(P, nk);
            std::vector<double> y(prods.size());
            double sum = 0.0;
            for (auto &v : y) { v = g.uni(0.02, 0.6); sum += v; }
            if (sum > 1.0) for (auto &v : y) v /= sum;
            for (size_t q = 0; q < prods.size(); ++q) {
                o->row_ind.push_back(r0 + prods[q]);
                o->vals.push_back(y[q]);
            }
            if (U > 0) {
                const int64_t unit = static_cast<int64_t>(g.below(static_cast<uint64_t>(U)));
                o->row_ind.push_back(r0 + P + Q + unit);
                o->vals.push_back(1.0);
            }
            o->c[static_cast<size_t>(col)] = g.uni(40.0, 90.0);
            o->u[static_cast<size_t>(col)] = g.uni(0.0, 1000.0);
            o->col_ptr[static_cast<size_t>(col + 1)] = static_cast<int64_t>(o->row_ind.size());
        }
        // sale columns: -1 in the product balance, blend coefficients
        for (int32_t k = 0; k < P; ++k) {
            const int64_t col = c0 + NC + k;
            o->row_ind.push_back(r0 + k);
            o->vals.push_back(-1.0);
            for (auto &qb : sale_blend[static_cast<size_t>(k)]) {
                o->row_ind.push_back(r0 + P + qb.first);
                o->vals.push_back(qb.second);
            }
            o->c[static_cast<size_t>(col)] = -g.uni(60.0, 120.0);
            o->u[static_cast<size_t>(col)] = g.uni(0.0, 1000.0);
            o->col_ptr[static_cast<size_t>(col + 1)] = static_cast<int64_t>(o->row_ind.size());
        }
        // inventory columns: leave block t, enter block t+1
        for (int32_t k = 0; k < P; ++k) {
            const int64_t col = c0 + NC + P + k;
            o->row_ind.push_back(r0 + k);
            o->vals.push_back(-1.0);
            if (t + 1 < T) {
                o->row_ind.push_back(r0 + R + k);
                o->vals.push_back(1.0);
            }
            o->c[static_cast<size_t>(col)] = g.uni(0.1, 1.0);
            o->u[static_cast<size_t>(col)] = g.uni(0.0, 1000.0);
            o->col_ptr[static_cast<size_t>(col + 1)] = static_cast<int64_t>(o->row_ind.size());
        }
    }
    finish_views(o);
    return o;
}

OwnedLp *gen_scenario(const OwnedLp *base, uint64_t seed, double shock, int32_t zero_cap) {
    auto *o = new OwnedLp(*base);
    Rng g(seed);
    for (auto &cj : o->c) cj *= std::exp(shock * g.normal());
    if (zero_cap && !o->row_kind.empty()) {
        std::vector<int64_t> caps;
        for (int64_t i = 0; i < o->m; ++i)
            if (o->row_kind[static_cast<size_t>(i)] == 2) caps.push_back(i);
        if (!caps.empty())
            o->ru[static_cast<size_t>(caps[g.below(caps.size())])] = 0.0;
    }
    finish_views(o);
    return o;
}

// The perturbations of py/refinery/parameterized_generator.py:28-133 on the
// staircase model: a crude column is a crude purchase (its cost the crude
// price, its upper bound the supply), a sale column a product sale (its cost
// minus the product price, its upper bound the demand cap), a capacity row a
// unit's capacity.  Each multiplier is 1 + clip(N(0, std), lo, hi) with the
// reference's clips; the regime factors follow :63-113.  The draws run in a
// fixed order (costs, then supplies, capacities, demands, then the outage),
// each over the LP in index order.
OwnedLp *gen_scenario_ex(const OwnedLp *base, uint64_t seed, const mpdhg_scenario_opts &op) {
    auto *o = new OwnedLp(*base);
    Rng g(seed);
    auto mult = [&](double sd, double lo, double hi) {
        if (sd <= 0.0) return 1.0;
        return 1.0 + std::max(lo, std::min(hi, sd * g.normal()));
    };
    if (op.shock_sigma > 0.0)
        for (auto &cj : o->c) cj *= std::exp(op.shock_sigma * g.normal());
    const bool stair = !o->col_kind.empty();
    if (stair) {
        // position of a crude / sale column within its block
        std::vector<int32_t> pos(static_cast<size_t>(o->n), 0);
        for (int64_t j = 0, run = 0; j < o->n; ++j) {
            const bool same = j > 0 && o->col_kind[static_cast<size_t>(j)] == o->col_kind[static_cast<size_t>(j - 1)] &&
                              o->col_block[static_cast<size_t>(j)] == o->col_block[static_cast<size_t>(j - 1)];
            run = same ? run + 1 : 0;
            pos[static_cast<size_t>(j)] = static_cast<int32_t>(run);
        }
        const int32_t rg = op.regime;
        for (int64_t j = 0; j < o->n; ++j) {  // prices
            const int8_t k = o->col_kind[static_cast<size_t>(j)];
            const int32_t p = pos[static_cast<size_t>(j)];
            double f = 1.0;
            if (k == 0) {
                f = mult(op.price_std, -0.30, 0.40);
                if (rg == MPDHG_REGIME_HIGH_CRUDE) f *= 1.30;
                if (rg == MPDHG_REGIME_SOUR_CRUDE) f *= (p == 0 ? 0.85 : (p == 1 ? 1.25 : 1.10));
            } else if (k == 1) {
                f = mult(op.price_std, -0.25, 0.35);
                if (rg == MPDHG_REGIME_DIESEL_SURGE) f *= (p % 4 == 0) ? 1.30 * 1.25 : 1.30;
            }
            o->c[static_cast<size_t>(j)] *= f;
        }
        for (int64_t j = 0; j < o->n; ++j) {  // crude supplies
            if (o->col_kind[static_cast<size_t>(j)] != 0) continue;
            double f = mult(op.supply_std, -0.40, 0.50);
            if (rg == MPDHG_REGIME_SWEET_SHORTAGE && pos[static_cast<size_t>(j)] == 1) f *= 0.35;
            o->u[static_cast<size_t>(j)] *= f;
        }
        std::vector<int32_t> cap_pos(static_cast<size_t>(o->m), 0);
        for (int64_t i = 0, run = 0; i < o->m; ++i) {
            const bool same = i > 0 && o->row_kind[static_cast<size_t>(i)] == o->row_kind[static_cast<size_t>(i - 1)] &&
                              o->row_block[static_cast<size_t>(i)] == o->row_block[static_cast<size_t>(i - 1)];
            run = same ? run + 1 : 0;
            cap_pos[static_cast<size_t>(i)] = static_cast<int32_t>(run);
        }
        for (int64_t i = 0; i < o->m; ++i) {  // unit capacities
            if (o->row_kind[static_cast<size_t>(i)] != 2) continue;
            double f = mult(op.capacity_std, -0.25, 0.30);
            if (rg == MPDHG_REGIME_TURNAROUND && cap_pos[static_cast<size_t>(i)] == 0) f *= 0.65;
            o->ru[static_cast<size_t>(i)] *= f;
        }
        for (int64_t j = 0; j < o->n; ++j) {  // demand caps
            if (o->col_kind[static_cast<size_t>(j)] != 1) continue;
            double f = mult(op.demand_std, -0.20, 0.25);
            if (rg == MPDHG_REGIME_DIESEL_SURGE && pos[static_cast<size_t>(j)] % 4 == 0) f *= 1.60 / 1.25;
            o->u[static_cast<size_t>(j)] *= f;
        }
    }
    if (op.zero_capacity && !o->row_kind.empty()) {
        std::vector<int64_t> caps;
        for (int64_t i = 0; i < o->m; ++i)
            if (o->row_kind[static_cast<size_t>(i)] == 2) caps.push_back(i);
        if (!caps.empty()) o->ru[static_cast<size_t>(caps[g.below(caps.size())])] = 0.0;
    }
    finish_views(o);
    return o;
}

}  // namespace mpdhg

extern "C" {

void mpdhg_scenario_opts_default(mpdhg_scenario_opts *opts) {
    if (opts) std::memset(opts, 0, sizeof(*opts));
}

int mpdhg_gen_scenario_ex(const mpdhg_lp_owned *base, uint64_t seed, const mpdhg_scenario_opts *opts,
                          mpdhg_lp_owned **out) {
    if (!base || !opts || !out) return mpdhg::fail(MPDHG_E_ARG, "gen_scenario_ex: NULL argument");
    if (opts->regime < MPDHG_REGIME_NORMAL || opts->regime > MPDHG_REGIME_SWEET_SHORTAGE)
        return mpdhg::fail(MPDHG_E_ARG, "gen_scenario_ex: unknown regime");
    if (opts->shock_sigma < 0 || opts->price_std < 0 || opts->supply_std < 0 || opts->capacity_std < 0 ||
        opts->demand_std < 0)
        return mpdhg::fail(MPDHG_E_ARG, "gen_scenario_ex: negative standard deviation");
    try {
        *out = reinterpret_cast<mpdhg_lp_owned *>(
            mpdhg::gen_scenario_ex(reinterpret_cast<const mpdhg::OwnedLp *>(base), seed, *opts));
    } catch (const std::bad_alloc &) {
        return mpdhg::fail(MPDHG_E_NOMEM, "gen_scenario_ex: out of host memory");
    }
    return MPDHG_OK;
}

int mpdhg_gen_random(int64_t m, int64_t n, int32_t k, uint64_t seed, mpdhg_lp_owned **out) {
    if (!out || m < 1 || n < 1 || k < 1) return mpdhg::fail(MPDHG_E_ARG, "gen_random: bad size");
    try {
        *out = reinterpret_cast<mpdhg_lp_owned *>(mpdhg::gen_random(m, n, k, seed));
    } catch (const std::bad_alloc &) {
        return mpdhg::fail(MPDHG_E_NOMEM, "gen_random: out of host memory");
    }
    return MPDHG_OK;
}

int mpdhg_gen_staircase(int32_t T, int32_t R, int32_t C, uint64_t seed, mpdhg_lp_owned **out) {
    if (!out) return mpdhg::fail(MPDHG_E_ARG, "gen_staircase: out is NULL");
    mpdhg::OwnedLp *o = nullptr;
    try {
        o = mpdhg::gen_staircase(T, R, C, seed);
    } catch (const std::bad_alloc &) {
        return mpdhg::fail(MPDHG_E_NOMEM, "gen_staircase: out of host memory");
    }
    if (!o) return mpdhg::fail(MPDHG_E_ARG, "gen_staircase: need T>=1, R>=10, C>2*(R/2)");
    *out = reinterpret_cast<mpdhg_lp_owned *>(o);
    return MPDHG_OK;
}

int mpdhg_gen_scenario(const mpdhg_lp_owned *base, uint64_t seed, double shock, int32_t zero_cap,
                       mpdhg_lp_owned **out) {
    if (!base || !out) return mpdhg::fail(MPDHG_E_ARG, "gen_scenario: NULL argument");
    try {
        *out = reinterpret_cast<mpdhg_lp_owned *>(
            mpdhg::gen_scenario(reinterpret_cast<const mpdhg::OwnedLp *>(base), seed, shock, zero_cap));
    } catch (const std::bad_alloc &) {
        return mpdhg::fail(MPDHG_E_NOMEM, "gen_scenario: out of host memory");
    }
    return MPDHG_OK;
}

const mpdhg_lp *mpdhg_lp_view(const mpdhg_lp_owned *lp) {
    return lp ? &reinterpret_cast<const mpdhg::OwnedLp *>(lp)->view : nullptr;
}

const int8_t *mpdhg_lp_row_kind(const mpdhg_lp_owned *lp) {
    if (!lp) return nullptr;
    auto *o = reinterpret_cast<const mpdhg::OwnedLp *>(lp);
    return o->row_kind.empty() ? nullptr : o->row_kind.data();
}

void mpdhg_lp_free(mpdhg_lp_owned *lp) { delete reinterpret_cast<mpdhg::OwnedLp *>(lp); }

}  // extern "C"
