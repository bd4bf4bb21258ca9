// extern "C" entry points of libmpdhg (declared in include/mpdhg.h).
// Every call catches internal exceptions and turns them into an MPDHG_E_*
// code plus a thread-local message.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "engine.hpp"
#include "mpdhg.h"
#include "mpdhg_internal.hpp"

namespace mpdhg {
namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

}  // namespace mpdhg

struct mpdhg_ctx {
    mpdhg::Context c;
};

namespace {

template <typename F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const mpdhg::Error &e) {
        return mpdhg::fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return mpdhg::fail(MPDHG_E_NOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return mpdhg::fail(MPDHG_E_STATE, e.what());
    }
}

int check_config(const mpdhg_config *cfg) {
    if (!cfg) return mpdhg::fail(MPDHG_E_ARG, "config is NULL");
    if (cfg->mode != MPDHG_MODE_REFERENCE && cfg->mode != MPDHG_MODE_ENHANCED)
        return mpdhg::fail(MPDHG_E_ARG, "unknown mode");
    if (cfg->check_frequency <= 0) return mpdhg::fail(MPDHG_E_ARG, "check_frequency must be > 0");
    if (cfg->spmv_layout != 0 && cfg->spmv_layout != 1)
        return mpdhg::fail(MPDHG_E_ARG, "spmv_layout must be 0 (auto) or 1 (storage order)");
    if (cfg->spmv_chunk != 0 && !mpdhg::valid_chunk(cfg->spmv_chunk))
        return mpdhg::fail(MPDHG_E_ARG, "spmv_chunk must be 0 (auto), 1024, 2048, 4096 or 8192");
    if (cfg->spmv_block != 0 && !mpdhg::valid_block(cfg->spmv_block))
        return mpdhg::fail(MPDHG_E_ARG, "spmv_block must be 0 (auto), 128, 256 or 512");
    if (cfg->spmv_block != 0 && cfg->spmv_chunk != 0 &&
        !mpdhg::valid_schedule(cfg->spmv_block, cfg->spmv_chunk))
        return mpdhg::fail(MPDHG_E_ARG, "(spmv_block, spmv_chunk) is not a compiled tile schedule");
    if (cfg->spmv_tune < 0 || cfg->spmv_tune > 2)
        return mpdhg::fail(MPDHG_E_ARG, "spmv_tune must be 0, 1 or 2");
    if (cfg->launch_graph != 0 && cfg->launch_graph != 1)
        return mpdhg::fail(MPDHG_E_ARG, "launch_graph must be 0 or 1");
    if (cfg->cpu_fallback != 0 && cfg->cpu_fallback != 1)
        return mpdhg::fail(MPDHG_E_ARG, "cpu_fallback must be 0 or 1");
    if (cfg->spmv_bands < 0 || cfg->spmv_bands > mpdhg::kMaxBands)
        return mpdhg::fail(MPDHG_E_ARG, "spmv_bands must be in [0, 16]");
    if (cfg->vertex_check < 0) return mpdhg::fail(MPDHG_E_ARG, "vertex_check must be >= 0");
    return MPDHG_OK;
}

int make_ctx(int device, mpdhg_ctx **out) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return mpdhg::fail(MPDHG_E_NODEV, "no HIP device visible");
    if (device < 0 || device >= n) return mpdhg::fail(MPDHG_E_ARG, "device index out of range");
    return guarded([&] {
        MPDHG_HIP(hipSetDevice(device));
        mpdhg::noalloc_setup(device);  // scratch of the no-allocation solve, once per device
        auto *c = new mpdhg_ctx();
        c->c.device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            mpdhg::hip_check(e, "hipStreamCreateWithFlags");
        }
        *out = c;
        return MPDHG_OK;
    });
}

int visible_devices() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    return n;
}

// GpuSolver::solve_pdhg without a device (gpu_solver.cpp:66-71): a notice on
// stderr, then the host FP64 simplex on each LP (mpdhg_polish from the slack
// basis).  The simplex's duals carry the simplex sign (reduced cost =
// c - A^T y_s), so the PDHG-sign duals are -y_s and the reduced costs are the
// same in both conventions.
int host_fallback(const mpdhg_lp *lps, int count, const mpdhg_config *cfg, mpdhg_result *res) {
    std::fprintf(stderr, "[mpdhg] No HIP device found. Falling back to the CPU simplex.\n");
    mpdhg_polish_opts po;
    mpdhg_polish_opts_default(&po);
    for (int q = 0; q < count; ++q) {
        const double t0 = mpdhg::now_sec();
        const mpdhg_lp &lp = lps[q];
        mpdhg_result &r = res[q];
        if (lp.m < 0 || lp.n < 0) return mpdhg::fail(MPDHG_E_ARG, "malformed LP");
        const size_t m = static_cast<size_t>(lp.m), n = static_cast<size_t>(lp.n);
        std::vector<double> x(n), ys(m), rc(n), ax(m);
        mpdhg_polish_result pr;
        std::memset(&pr, 0, sizeof(pr));
        pr.x = x.data();
        pr.y = ys.data();
        pr.reduced_costs = rc.data();
        pr.row_activity = ax.data();
        const int rc_code = mpdhg_polish(&lp, nullptr, nullptr, nullptr, &po, &pr);
        if (rc_code != MPDHG_OK) return rc_code;
        int rc2 = guarded([&] {
            mpdhg::HostLp h;
            mpdhg::prepare_structure(lp, h);
            std::vector<double> y(m);
            for (size_t i = 0; i < m; ++i) y[i] = -ys[i];
            mpdhg::host_kkt(lp, h, x, y, r);
            const double sg = cfg->dual_sign ? -1.0 : 1.0;
            if (r.x) std::memcpy(r.x, x.data(), n * sizeof(double));
            if (r.y)
                for (size_t i = 0; i < m; ++i) r.y[i] = sg * y[i];
            return MPDHG_OK;
        });
        if (rc2) return rc2;
        r.status = pr.status;
        r.iterations = pr.iterations;
        r.primal_objective = pr.objective;
        r.time_wall_sec = mpdhg::now_sec() - t0;
        r.solve_sec = pr.time_sec;
    }
    return MPDHG_OK;
}

void clear_result(mpdhg_result *r) {
    double *x = r->x, *y = r->y, *rc = r->reduced_costs, *ra = r->row_activity;
    std::memset(r, 0, sizeof(*r));
    r->x = x;
    r->y = y;
    r->reduced_costs = rc;
    r->row_activity = ra;
    r->status = MPDHG_STATUS_UNKNOWN;
    r->csr_bands = 1;
    r->vertex_status = MPDHG_STATUS_UNKNOWN;
    r->vertex_objective = std::numeric_limits<double>::quiet_NaN();
    r->vertex_rel_error = std::numeric_limits<double>::quiet_NaN();
}

// The vertex check needs each LP's point even where the caller skips x:
// such results get a temporary x for the solve.
struct XHold {
    std::vector<std::vector<double>> tmp;
    std::vector<const double *> x;
    XHold(const mpdhg_lp *lps, int count, mpdhg_result *res) {
        tmp.resize(static_cast<size_t>(count));
        for (int q = 0; q < count; ++q) {
            if (!res[q].x && lps[q].n > 0) {
                tmp[static_cast<size_t>(q)].resize(static_cast<size_t>(lps[q].n));
                res[q].x = tmp[static_cast<size_t>(q)].data();
            }
            x.push_back(res[q].x);
        }
    }
    void restore(mpdhg_result *res) {
        for (size_t q = 0; q < tmp.size(); ++q)
            if (!tmp[q].empty()) res[q].x = nullptr;
    }
};

// mpdhg_solve_batch / mpdhg_solve_warm: the device solve, then the optional
// vertex check.
int run_solve(mpdhg_ctx *ctx, const mpdhg_lp *lps, int count, const mpdhg_config *cfg,
              const mpdhg_warm_start *warm, mpdhg_result *res) {
    if (count < 0 || (count > 0 && (!lps || !res))) return mpdhg::fail(MPDHG_E_ARG, "bad batch");
    int rc = check_config(cfg);
    if (rc) return rc;
    for (int q = 0; q < count; ++q) clear_result(&res[q]);
    if (!ctx && cfg->cpu_fallback && visible_devices() <= 0) return host_fallback(lps, count, cfg, res);
    mpdhg_ctx *tmp = nullptr;
    if (!ctx) {
        rc = make_ctx(0, &tmp);
        if (rc) return rc;
        ctx = tmp;
    }
    rc = guarded([&] {
        MPDHG_HIP(hipSetDevice(ctx->c.device));
        XHold hold(lps, count, res);
        try {
            if (cfg->mode == MPDHG_MODE_REFERENCE)
                mpdhg::solve_reference(ctx->c, lps, count, *cfg, warm, res);
            else
                mpdhg::solve_enhanced(ctx->c, lps, count, *cfg, warm, res);
            if (cfg->vertex_check > 0) mpdhg::vertex_check(lps, count, hold.x, cfg->vertex_check, res);
        } catch (...) {
            hold.restore(res);
            throw;
        }
        hold.restore(res);
        return MPDHG_OK;
    });
    delete tmp;
    return rc;
}

}  // namespace

extern "C" {

const char *mpdhg_last_error(void) { return mpdhg::g_last_error.c_str(); }

void mpdhg_config_default(mpdhg_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->max_iterations = 4000;  // gpu_solver.hpp:13-21
    cfg->tol_primal = 1e-4f;
    cfg->tol_dual = 1e-4f;
    cfg->tol_gap = 1e-4f;
    cfg->check_frequency = 25;
    cfg->step_tau = 0.0f;
    cfg->step_sigma = 0.0f;
    cfg->enable_restart = 1;
    cfg->enable_hybrid_polish = 1;
    cfg->mode = MPDHG_MODE_REFERENCE;
    cfg->tol_kkt = 0.0;
    cfg->max_refinements = 0;
    cfg->reflection = 1.0f;
    cfg->ruiz_iterations = 10;
    cfg->profile = 0;
    cfg->dual_sign = 0;
    cfg->refine_ratio = 1e-2;
    cfg->spmv_layout = 0;
    cfg->spmv_chunk = 0;
    cfg->launch_graph = 1;
    cfg->spmv_block = 0;
    cfg->spmv_tune = 1;
}

int mpdhg_device_count(int *count) {
    if (!count) return mpdhg::fail(MPDHG_E_ARG, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return MPDHG_OK;
}

int mpdhg_mem_info(int device, size_t *total, size_t *free_b, size_t *used) {
    if (!total || !free_b || !used) return mpdhg::fail(MPDHG_E_ARG, "NULL output");
    *total = *free_b = *used = 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return mpdhg::fail(MPDHG_E_NODEV, "no HIP device visible");
    if (device < 0 || device >= n) return mpdhg::fail(MPDHG_E_ARG, "device index out of range");
    return guarded([&] {
        MPDHG_HIP(hipSetDevice(device));
        size_t f = 0, t = 0;
        MPDHG_HIP(hipMemGetInfo(&f, &t));
        *total = t;
        *free_b = f;
        *used = t - f;
        return MPDHG_OK;
    });
}

int mpdhg_create(int device, mpdhg_ctx **out) {
    if (!out) return mpdhg::fail(MPDHG_E_ARG, "out is NULL");
    *out = nullptr;
    return make_ctx(device, out);
}

void mpdhg_destroy(mpdhg_ctx *ctx) { delete ctx; }

int mpdhg_solve_batch(mpdhg_ctx *ctx, const mpdhg_lp *lps, int count, const mpdhg_config *cfg,
                      mpdhg_result *res) {
    return run_solve(ctx, lps, count, cfg, nullptr, res);
}

int mpdhg_solve_warm(mpdhg_ctx *ctx, const mpdhg_lp *lps, int count, const mpdhg_config *cfg,
                     const mpdhg_warm_start *warm, mpdhg_result *res) {
    return run_solve(ctx, lps, count, cfg, warm, res);
}

int mpdhg_adjoint_batch(mpdhg_ctx *ctx, const mpdhg_lp *lps, int count, mpdhg_adjoint_io *io,
                        double active_tol, double tol, int max_iter) {
    if (count < 0 || (count > 0 && (!lps || !io))) return mpdhg::fail(MPDHG_E_ARG, "bad batch");
    mpdhg_ctx *tmp = nullptr;
    int rc = MPDHG_OK;
    if (!ctx) {
        rc = make_ctx(0, &tmp);
        if (rc) return rc;
        ctx = tmp;
    }
    rc = guarded([&] {
        MPDHG_HIP(hipSetDevice(ctx->c.device));
        mpdhg::solve_adjoint(ctx->c, lps, count, io, active_tol, tol, max_iter);
        return MPDHG_OK;
    });
    delete tmp;
    return rc;
}

int mpdhg_solve(mpdhg_ctx *ctx, const mpdhg_lp *lp, const mpdhg_config *cfg, mpdhg_result *res) {
    if (!lp || !res) return mpdhg::fail(MPDHG_E_ARG, "NULL lp or result");
    return mpdhg_solve_batch(ctx, lp, 1, cfg, res);
}

int mpdhg_solve_noalloc(mpdhg_ctx *ctx, const mpdhg_lp *lp, const mpdhg_config *cfg, mpdhg_result *res) {
    if (!lp || !res) return mpdhg::fail(MPDHG_E_ARG, "NULL lp or result");
    int rc = check_config(cfg);
    if (rc) return rc;
    if (lp->m < 0 || lp->n < 0 || lp->nnz < 0) return mpdhg::fail(MPDHG_E_ARG, "malformed LP");
    const size_t need = mpdhg::noalloc_bytes(cfg->mode, lp->m, lp->n, lp->nnz, cfg->spmv_layout == 0);
    if (need > MPDHG_NOALLOC_SCRATCH_BYTES)
        return mpdhg::fail(MPDHG_E_ARG, "LP too large for the no-allocation solve: up to " +
                                            std::to_string(need) + " bytes of scratch needed, " +
                                            std::to_string(MPDHG_NOALLOC_SCRATCH_BYTES) + " held");
    if (!ctx) return mpdhg::fail(MPDHG_E_ARG, "the no-allocation solve needs a context (mpdhg_create)");
    clear_result(res);
    mpdhg_config c = *cfg;
    c.profile = 0;
    c.launch_graph = 0;
    c.spmv_tune = 0;
    c.spmv_bands = 1;
    c.vertex_check = 0;
    return guarded([&] {
        MPDHG_HIP(hipSetDevice(ctx->c.device));
        mpdhg::Arena arena;
        arena.base = static_cast<char *>(mpdhg::noalloc_scratch(ctx->c.device));
        arena.cap = MPDHG_NOALLOC_SCRATCH_BYTES;
        mpdhg::ArenaScope scope(&arena);
        if (c.mode == MPDHG_MODE_REFERENCE)
            mpdhg::solve_reference(ctx->c, lp, 1, c, nullptr, res);
        else
            mpdhg::solve_enhanced(ctx->c, lp, 1, c, nullptr, res);
        return MPDHG_OK;
    });
}

}  // extern "C"
