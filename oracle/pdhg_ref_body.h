/*
 * oracle/pdhg_ref_body.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Body of the reference-mode PDHG restatement, instantiated twice by
 * pdhg_oracle.c: once with R = float (the reference device arithmetic) and
 * once with R = double (the FP64 CPU path of the CPU-runnable workload).
 * Included with these macros defined:
 *   R        element type (float | double)
 *   RFMA     fused multiply-add for R (fmaf | fma)
 *   RABS     absolute value for R
 *   RSFX     name suffix (f32 | f64)
 *   R_BIG    the reference's "infinite bound" cut (1e15f, 1e15)
 *   R_TAUMIN the residual step floor (1e-4f)
 *
 * Every line follows core/src/gpu/gpu_solver.cpp:60-307 (host algorithm)
 * and core/src/gpu/cuda_pdhg.cu:14-151 (kernel arithmetic).  Where nvcc's
 * default -fmad=true contracts `a + b*c`/`a - b*c` into one FMA, the
 * restatement writes RFMA explicitly; the file is compiled with
 * -ffp-contract=off so nothing else is fused.
 */

#define ORC_CAT2(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC_NAME(base) ORC_CAT(base, RSFX)

/* cuda_pdhg.cu:14-30 csr_spmv_kernel (and :33-49 for the CSC form):
 * sum over the row in storage order, sum = fma(v, x[col], sum). */
static void ORC_NAME(seq_spmv)(int rows, const int *ptr, const int *idx,
                               const R *val, const R *xs, R *out) {
    for (int r = 0; r < rows; ++r) {
        R sum = (R)0;
        for (int k = ptr[r]; k < ptr[r + 1]; ++k) sum = RFMA(val[k], xs[idx[k]], sum);
        out[r] = sum;
    }
}

/* std::max / std::min as the reference host code calls them. */
static inline R ORC_NAME(smax)(R a, R b) { return (a < b) ? b : a; }
static inline R ORC_NAME(smin)(R a, R b) { return (b < a) ? b : a; }

/* gpu_solver.cpp:112-151 -- normalisers and diagonal steps. */
void ORC_NAME(orc_ref_setup)(int m, int n, const int *col_ptr, const double *vals_d,
                             const int *row_ptr, const R *csr_vals,
                             const double *c, const double *rl, const double *ru,
                             R step_tau, R step_sigma,
                             R *tau, R *sigma, R *max_abs_c, R *max_abs_b) {
    R mac = (R)1;
    for (int j = 0; j < n; ++j) mac = ORC_NAME(smax)(mac, RABS((R)c[j]));
    R mab = (R)1;
    for (int i = 0; i < m; ++i) {
        R hru = (R)ru[i], hrl = (R)rl[i];
        if (hru < R_BIG) mab = ORC_NAME(smax)(mab, RABS(hru));
        if (hrl > -R_BIG) mab = ORC_NAME(smax)(mab, RABS(hrl));
    }
    for (int j = 0; j < n; ++j) {
        R s = (R)0;
        for (int k = col_ptr[j]; k < col_ptr[j + 1]; ++k) s = s + RABS((R)vals_d[k]);
        tau[j] = (step_tau > (R)0) ? step_tau : (R)0.95 / ORC_NAME(smax)((R)1, s);
    }
    for (int i = 0; i < m; ++i) {
        R s = (R)0;
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) s = s + RABS(csr_vals[k]);
        sigma[i] = (step_sigma > (R)0) ? step_sigma : (R)0.95 / ORC_NAME(smax)((R)1, s);
    }
    *max_abs_c = mac;
    *max_abs_b = mab;
}

/* gpu_solver.cpp:60-307 + cuda_pdhg.cu:14-151, with a correct A^T y
 * (the reference device uploads int64 CSC indices with int32 byte counts,
 * gpu_solver.cpp:202-203; that defect is not restated here). */
int ORC_NAME(orc_pdhg_ref)(int m, int n, const int *col_ptr, const int *row_ind,
                           const double *vals_d, const double *c, const double *l,
                           const double *u, const double *rl, const double *ru,
                           const orc_ref_cfg *cfg, double *x_out, double *y_out,
                           R *x_last, R *xbar_last, orc_ref_out *out) {
    memset(out, 0, sizeof(*out));
    out->status = 8; /* Unknown */
    if (m == 0 || n == 0) { out->status = 0; return 0; }
    const int nnz = col_ptr[n];

    R *csc_v = (R *)malloc(sizeof(R) * (size_t)(nnz ? nnz : 1));
    int *rp = (int *)malloc(sizeof(int) * (size_t)(m + 1));
    int *ci = (int *)malloc(sizeof(int) * (size_t)(nnz ? nnz : 1));
    R *csr_v = (R *)calloc((size_t)(nnz ? nnz : 1), sizeof(R));
    R *hc = (R *)malloc(sizeof(R) * (size_t)n), *hl = (R *)malloc(sizeof(R) * (size_t)n);
    R *hu = (R *)malloc(sizeof(R) * (size_t)n), *tau = (R *)malloc(sizeof(R) * (size_t)n);
    R *hrl = (R *)malloc(sizeof(R) * (size_t)m), *hru = (R *)malloc(sizeof(R) * (size_t)m);
    R *sigma = (R *)malloc(sizeof(R) * (size_t)m);
    R *x = (R *)malloc(sizeof(R) * (size_t)n), *xbar = (R *)malloc(sizeof(R) * (size_t)n);
    R *y = (R *)calloc((size_t)m, sizeof(R));
    R *aty = (R *)malloc(sizeof(R) * (size_t)n), *axb = (R *)malloc(sizeof(R) * (size_t)m);
    R *ax = (R *)malloc(sizeof(R) * (size_t)m);
    R *best = (R *)malloc(sizeof(R) * (size_t)n);

    /* :88-89 value cast; :92-109 CSR by counting sort */
    for (int k = 0; k < nnz; ++k) csc_v[k] = (R)vals_d[k];
    memset(rp, 0, sizeof(int) * (size_t)(m + 1));
    for (int k = 0; k < nnz; ++k) if (row_ind[k] < m) rp[row_ind[k] + 1]++;
    for (int i = 0; i < m; ++i) rp[i + 1] += rp[i];
    {
        int *fill = (int *)malloc(sizeof(int) * (size_t)(m + 1));
        memcpy(fill, rp, sizeof(int) * (size_t)(m + 1));
        for (int j = 0; j < n; ++j)
            for (int k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
                int dst = fill[row_ind[k]]++;
                ci[dst] = j;
                csr_v[dst] = csc_v[k];
            }
        free(fill);
    }
    /* :111-126 FP-cast vectors */
    for (int j = 0; j < n; ++j) { hc[j] = (R)c[j]; hl[j] = (R)l[j]; hu[j] = (R)u[j]; }
    for (int i = 0; i < m; ++i) { hrl[i] = (R)rl[i]; hru[i] = (R)ru[i]; }
    R mac, mab;
    ORC_NAME(orc_ref_setup)(m, n, col_ptr, vals_d, rp, csr_v, c, rl, ru,
                            (R)cfg->step_tau, (R)cfg->step_sigma, tau, sigma, &mac, &mab);
    /* :192-196 initial point (or the optional start point, clamped alike) */
    for (int j = 0; j < n; ++j) {
        const R x0 = cfg->x0 ? (R)cfg->x0[j] : (R)0;
        x[j] = ORC_NAME(smax)(hl[j], ORC_NAME(smin)(hu[j], x0));
        xbar[j] = x[j];
        best[j] = x[j];
    }
    if (cfg->y0)
        for (int i = 0; i < m; ++i) y[i] = (R)cfg->y0[i];
    R best_res = (R)1e30;
    int iter, converged = 0;
    const int K = cfg->max_iterations, F = cfg->check_frequency;
    for (iter = 1; iter <= K; ++iter) {
        /* :229-230 A^T y then primal update (cuda_pdhg.cu:54-77) */
        ORC_NAME(seq_spmv)(n, col_ptr, row_ind, csc_v, y, aty);
        for (int j = 0; j < n; ++j) {
            R xo = x[j];
            R g = hc[j] + aty[j];
            R xn = RFMA(-tau[j], g, xo);
            if (xn < hl[j]) xn = hl[j];
            if (xn > hu[j]) xn = hu[j];
            x[j] = xn;
            xbar[j] = (R)2 * xn - xo;
        }
        /* :233-234 A xbar then dual update (cuda_pdhg.cu:82-103) */
        ORC_NAME(seq_spmv)(m, rp, ci, csr_v, xbar, axb);
        for (int i = 0; i < m; ++i) {
            R sg = sigma[i];
            R w = axb[i] + (y[i] / sg);
            R p = w;
            if (p < hrl[i]) p = hrl[i];
            if (p > hru[i]) p = hru[i];
            y[i] = sg * (w - p);
        }
        /* :237-239 restart kernel x̄ = x (cuda_pdhg.cu:106-111) */
        if (cfg->enable_restart && (iter % 250 == 0))
            for (int j = 0; j < n; ++j) xbar[j] = x[j];
        /* :241-270 convergence check */
        if (iter % F == 0 || iter == K) {
            ORC_NAME(seq_spmv)(m, rp, ci, csr_v, x, ax);
            ORC_NAME(seq_spmv)(n, col_ptr, row_ind, csc_v, y, aty);
            R mp = (R)0, md = (R)0;
            for (int i = 0; i < m; ++i) { /* cuda_pdhg.cu:114-129 */
                R a = ax[i], v = (R)0;
                if (hrl[i] > -R_BIG && a < hrl[i]) v = RMAXF(v, hrl[i] - a);
                if (hru[i] < R_BIG && a > hru[i]) v = RMAXF(v, a - hru[i]);
                mp = ORC_NAME(smax)(mp, v);
            }
            for (int j = 0; j < n; ++j) { /* cuda_pdhg.cu:131-151 */
                R xj = x[j];
                R g = hc[j] + aty[j];
                R pr = RFMA(-tau[j], g, xj);
                if (pr < hl[j]) pr = hl[j];
                if (pr > hu[j]) pr = hu[j];
                R v = RABS(xj - pr) / RMAXF(R_TAUMIN, tau[j]);
                md = ORC_NAME(smax)(md, v);
            }
            R P = mp / mab, D = md / mac;
            R comb = ORC_NAME(smax)(P, D);
            out->checks++;
            out->last_primal = (double)P;
            out->last_dual = (double)D;
            if (comb < best_res) {
                best_res = comb;
                memcpy(best, x, sizeof(R) * (size_t)n);
            }
            if (P <= (R)cfg->tol_primal && D <= (R)cfg->tol_dual) { converged = 1; break; }
        }
    }
    out->status = converged ? 0 : 4; /* Optimal : IterationLimit */
    out->iterations = iter;
    out->best_residual = (double)best_res;
    out->max_abs_c = (double)mac;
    out->max_abs_b = (double)mab;
    double obj = 0.0;
    for (int j = 0; j < n; ++j) {
        double xv = (double)best[j];
        if (x_out) x_out[j] = xv;
        obj += c[j] * xv;
    }
    out->objective = obj;
    if (y_out) for (int i = 0; i < m; ++i) y_out[i] = (double)y[i];
    if (x_last) memcpy(x_last, x, sizeof(R) * (size_t)n);
    if (xbar_last) memcpy(xbar_last, xbar, sizeof(R) * (size_t)n);

    free(csc_v); free(rp); free(ci); free(csr_v); free(hc); free(hl); free(hu); free(tau);
    free(hrl); free(hru); free(sigma); free(x); free(xbar); free(y); free(aty); free(axb);
    free(ax); free(best);
    return 0;
}

#undef ORC_NAME
#undef ORC_CAT
#undef ORC_CAT2
