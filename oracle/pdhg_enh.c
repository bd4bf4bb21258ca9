/*
 * oracle/pdhg_enh.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU port of the engine's enhanced mode: the same algorithm the GPU runs,
 * written as plain loops so that tests can compare iterates and the bench
 * can time a CPU baseline of the same method.  No reference counterpart:
 * the reference's solver has fixed diagonal steps and no restarts
 * (gpu_solver.cpp:128-151, :237-239).  Published methods:
 *   - Ruiz equilibration + Pock-Chambolle (alpha = 1) diagonal scaling
 *     (Pock & Chambolle, ICCV 2011; Applegate et al., NeurIPS 2021);
 *   - reflected restarted Halpern PDHG with fixed-point-residual restarts
 *     (Lu & Yang, "Restarted Halpern PDHG for linear programming", 2024);
 *   - primal weight at restarts by a PI controller on log(omega |dx|/|dy|)
 *     (Lu & Yang, "cuPDLPx", 2025; gains 0.6 and 0.05);
 *   - iterative refinement on the residual LP (Gleixner, Steffy & Wolter,
 *     INFORMS J. Comput. 2016), FP32 inner solves with FP64 outer updates,
 *     taken at restarts once the KKT error fell by refine_ratio.
 * FP32 iterate arithmetic follows the GPU kernels operation by operation
 * (explicit fmaf; compiled with -ffp-contract=off).  Every FP64 quantity
 * behind a decision (the power iteration's |z|^2, the check's KKT terms and
 * objectives, the fixed-point residuals and anchor distances) is formed in
 * the engine's fixed order: per-row (per-column) products as FMA chains in
 * storage order with the long-row split rule, 16-row group partials by a
 * lane butterfly, and per-instance sums over the group partials by the
 * engine's 256-lane tree (grp_part, canon).  So the whole solve, restarts,
 * refinements and stop included, is the GPU's bit for bit.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pdhg_oracle.h"

typedef struct {
    int m, n;
    const int *cp, *ri;
    const double *v;
    int *rp, *ci, *c2r;
    double *rv;               /* the FP64 values in CSR order */
    double *dr, *dc;          /* row / column scaling */
    float *cv_s, *rv_s;       /* scaled FP32 values, CSC and CSR order */
} enh_lp;

static void *xmalloc(size_t n) { return malloc(n ? n : 1); }

/* Ruiz (infinity-norm) passes, then one Pock-Chambolle (alpha = 1) pass. */
static void scale_lp(enh_lp *P, int ruiz) {
    const int m = P->m, n = P->n;
    double *rs = (double *)calloc((size_t)m + 1, sizeof(double));
    double *cs = (double *)calloc((size_t)n + 1, sizeof(double));
    for (int i = 0; i < m; ++i) P->dr[i] = 1.0;
    for (int j = 0; j < n; ++j) P->dc[j] = 1.0;
    for (int pass = 0; pass <= ruiz; ++pass) {
        const int l1 = (pass == ruiz);
        memset(rs, 0, sizeof(double) * (size_t)m);
        for (int j = 0; j < n; ++j) {
            double acc = 0.0;
            for (int k = P->cp[j]; k < P->cp[j + 1]; ++k) {
                const int i = P->ri[k];
                const double a = fabs(P->v[k]) * P->dr[i] * P->dc[j];
                if (l1) { acc += a; rs[i] += a; }
                else { if (a > acc) acc = a; if (a > rs[i]) rs[i] = a; }
            }
            cs[j] = acc;
        }
        for (int i = 0; i < m; ++i) if (rs[i] > 0.0) P->dr[i] /= sqrt(rs[i]);
        for (int j = 0; j < n; ++j) if (cs[j] > 0.0) P->dc[j] /= sqrt(cs[j]);
    }
    for (int j = 0; j < n; ++j)
        for (int k = P->cp[j]; k < P->cp[j + 1]; ++k) {
            const float s = (float)(P->dr[P->ri[k]] * P->v[k] * P->dc[j]);
            P->cv_s[k] = s;
            P->rv_s[P->c2r[k]] = s;
        }
    free(rs);
    free(cs);
}

/* FP32 row sums as the enhanced-mode tile kernels define them: a row with at
 * most ORC_SPLIT_ROW nonzeros is one sequential FMA chain from zero; a longer
 * row is summed in two levels counted from its first nonzero: segments of
 * ORC_SPLIT_SEG nonzeros, each a sequential FMA chain from zero; blocks of
 * ORC_SPLIT_BLOCK nonzeros, each the in-order sum of its segment sums; and
 * the row sum, the in-order sum of the block sums (kSplitRow, kSplitSeg and
 * kSplitBlockNnz of the engine). */
#define ORC_SPLIT_ROW 8192
#define ORC_SPLIT_SEG 64
#define ORC_SPLIT_BLOCK 4096
static void spmv32(int rows, const int *ptr, const int *idx, const float *v, const float *x,
                   float *out) {
    for (int r = 0; r < rows; ++r) {
        const int k0 = ptr[r], k1 = ptr[r + 1];
        float s = 0.0f;
        if (k1 - k0 <= ORC_SPLIT_ROW) {
            for (int k = k0; k < k1; ++k) s = fmaf(v[k], x[idx[k]], s);
        } else {
            for (int b0 = k0; b0 < k1; b0 += ORC_SPLIT_BLOCK) {
                const int b1 = (b0 + ORC_SPLIT_BLOCK < k1) ? b0 + ORC_SPLIT_BLOCK : k1;
                float blk = 0.0f;
                for (int g0 = b0; g0 < b1; g0 += ORC_SPLIT_SEG) {
                    const int g1 = (g0 + ORC_SPLIT_SEG < b1) ? g0 + ORC_SPLIT_SEG : b1;
                    float g = 0.0f;
                    for (int k = g0; k < g1; ++k) g = fmaf(v[k], x[idx[k]], g);
                    blk = blk + g;
                }
                s = s + blk;
            }
        }
        out[r] = s;
    }
}

void orc_spmv_enh_f32(int rows, const int *ptr, const int *idx, const float *vals,
                      const float *x, float *out) {
    spmv32(rows, ptr, idx, vals, x, out);
}

/* The same row sums in FP64 (the engine's FP64 check products), from init 0. */
static void spmv64(int rows, const int *ptr, const int *idx, const double *v, const double *x,
                   double *out) {
    for (int r = 0; r < rows; ++r) {
        const int k0 = ptr[r], k1 = ptr[r + 1];
        double s = 0.0;
        if (k1 - k0 <= ORC_SPLIT_ROW) {
            for (int k = k0; k < k1; ++k) s = fma(v[k], x[idx[k]], s);
        } else {
            for (int b0 = k0; b0 < k1; b0 += ORC_SPLIT_BLOCK) {
                const int b1 = (b0 + ORC_SPLIT_BLOCK < k1) ? b0 + ORC_SPLIT_BLOCK : k1;
                double blk = 0.0;
                for (int g0 = b0; g0 < b1; g0 += ORC_SPLIT_SEG) {
                    const int g1 = (g0 + ORC_SPLIT_SEG < b1) ? g0 + ORC_SPLIT_SEG : b1;
                    double g = 0.0;
                    for (int k = g0; k < g1; ++k) g = fma(v[k], x[idx[k]], g);
                    blk = blk + g;
                }
                s = s + blk;
            }
        }
        out[r] = s;
    }
}

/* ---- the engine's fixed-order FP64 reductions ----
 * A lane butterfly over w values (w a power of two <= 64): step o = w/2, ...,
 * 1 replaces v_l by v_l + v_(l xor o) (fmax for maxima); the result is lane
 * 0's, as a wave's xor-shuffle reduction leaves it. */
#define ORC_GROUP 16     /* rows per group (kGroup) */
#define ORC_LANES 256    /* threads of the per-instance kernels (kInstBlock) */
#define ORC_WAVE 64
static double bfly(const double *v0, int w, int is_max) {
    double v[ORC_WAVE], t[ORC_WAVE];
    memcpy(v, v0, sizeof(double) * (size_t)w);
    for (int o = w / 2; o > 0; o >>= 1) {
        for (int l = 0; l < w; ++l) t[l] = is_max ? fmax(v[l], v[l ^ o]) : v[l] + v[l ^ o];
        memcpy(v, t, sizeof(double) * (size_t)w);
    }
    return v[0];
}

/* Group partials of val[0, n): group g holds elements 16 g ... 16 g + 15 (0
 * past n), combined by the 16-lane butterfly. */
static void grp_part(const double *val, int n, int is_max, double *part) {
    for (int g = 0; g * ORC_GROUP < n; ++g) {
        double v[ORC_GROUP];
        for (int l = 0; l < ORC_GROUP; ++l) {
            const int k = g * ORC_GROUP + l;
            v[l] = k < n ? val[k] : 0.0;
        }
        part[g] = bfly(v, ORC_GROUP, is_max);
    }
}

/* Instance total of G group partials: lane l of 256 adds partials l, l + 256,
 * ... in order from 0; each wave of 64 lanes combines by the butterfly; the
 * four wave results are added in order. */
static double canon(const double *part, int G, int is_max) {
    double acc[ORC_LANES];
    for (int l = 0; l < ORC_LANES; ++l) {
        double a = 0.0;
        for (int g = l; g < G; g += ORC_LANES) a = is_max ? fmax(a, part[g]) : a + part[g];
        acc[l] = a;
    }
    double r = 0.0;
    for (int w = 0; w < ORC_LANES / ORC_WAVE; ++w) {
        const double s = bfly(acc + w * ORC_WAVE, ORC_WAVE, is_max);
        r = (w == 0) ? s : (is_max ? fmax(r, s) : r + s);
    }
    return r;
}

/* canon of the group partials of val[0, n); part holds ceil(n / 16) doubles. */
static double canon_of(const double *val, int n, int is_max, double *part) {
    grp_part(val, n, is_max, part);
    return canon(part, (n + ORC_GROUP - 1) / ORC_GROUP, is_max);
}

/* Power iteration on A^T A for sigma_max of the scaled matrix. */
static double power_sigma(const enh_lp *P, int iters) {
    const int m = P->m, n = P->n;
    float *v = (float *)xmalloc(sizeof(float) * (size_t)n);
    float *w = (float *)xmalloc(sizeof(float) * (size_t)m);
    float *z = (float *)xmalloc(sizeof(float) * (size_t)n);
    double nv = 0.0, s = 0.0;
    for (int j = 0; j < n; ++j) {
        v[j] = 1.0f + 0.5f * (float)((j * 2654435761u) % 1000u) / 1000.0f;
        nv += (double)v[j] * v[j];
    }
    double *z2 = (double *)xmalloc(sizeof(double) * (size_t)n);
    double *part = (double *)xmalloc(sizeof(double) * (size_t)(n / ORC_GROUP + 1));
    for (int it = 0; it < iters && nv > 0.0; ++it) {
        const float inv = (float)(1.0 / sqrt(nv));
        for (int j = 0; j < n; ++j) v[j] *= inv;
        spmv32(m, P->rp, P->ci, P->rv_s, v, w);
        spmv32(n, P->cp, P->ri, P->cv_s, w, z);
        for (int j = 0; j < n; ++j) z2[j] = (double)z[j] * (double)z[j];
        const double nz = canon_of(z2, n, 0, part);
        s = sqrt(sqrt(nz)); /* ||A^T A v|| with ||v|| = 1 -> sigma^2 */
        memcpy(v, z, sizeof(float) * (size_t)n);
        nv = nz;
    }
    free(v); free(w); free(z); free(z2); free(part);
    return s;
}

/* FP64 relative KKT of (x, y) on the original LP as the engine's check forms
 * it (out: eps_p, eps_d, eps_gap, kkt, primal and dual objective); also A x
 * and lambda = c + A^T y.  w: scratch of 5 max(m, n) doubles. */
static void kkt_full(const enh_lp *P, const double *c, const double *l, const double *u,
                     const double *rl, const double *ru, const double *x, const double *y,
                     double bnorm, double cnorm, double *ax, double *lam, double *out, double *w) {
    const int m = P->m, n = P->n, mx = m > n ? m : n;
    double *q0 = w, *q1 = w + mx, *q2 = w + 2 * (size_t)mx, *part = w + 3 * (size_t)mx;
    spmv64(m, P->rp, P->ci, P->rv, x, ax);
    /* rows: bound violation of A x, sign violation of y, dual objective terms */
    for (int i = 0; i < m; ++i) {
        const double lo = rl[i], hi = ru[i], axi = ax[i];
        double viol = 0.0;
        if (isfinite(lo) && axi < lo) viol = lo - axi;
        if (isfinite(hi) && axi > hi) viol = axi - hi;
        const double yv = y[i];
        double yi = yv, dobj = 0.0;
        if (!isfinite(hi) && yi > 0.0) yi = 0.0;
        if (!isfinite(lo) && yi < 0.0) yi = 0.0;
        if (yi > 0.0) dobj = -hi * yi;
        else if (yi < 0.0) dobj = -lo * yi;
        q0[i] = viol;
        q1[i] = fabs(yv - yi);
        q2[i] = dobj;
    }
    const double rv = canon_of(q0, m, 1, part), sv = canon_of(q1, m, 1, part);
    const double dobjr = canon_of(q2, m, 0, part);
    /* columns: lambda, its violation of the bound signs, bound violation of x,
     * dual and primal objective terms */
    spmv64(n, P->cp, P->ri, P->v, y, lam);
    double *q3 = w + 4 * (size_t)mx;
    for (int j = 0; j < n; ++j) {
        const double lo = l[j], hi = u[j], xj = x[j], cj = c[j];
        const double lm = cj + lam[j];
        lam[j] = lm;
        double p = lm, dobj = 0.0, cv = 0.0;
        if (!isfinite(lo) && p > 0.0) p = 0.0;
        if (!isfinite(hi) && p < 0.0) p = 0.0;
        if (p > 0.0) dobj = lo * p;
        else if (p < 0.0) dobj = hi * p;
        if (isfinite(lo) && xj < lo) cv = lo - xj;
        if (isfinite(hi) && xj > hi) cv = xj - hi;
        q0[j] = fabs(lm - p);
        q1[j] = cv;
        q2[j] = cj * xj;
        q3[j] = dobj;
    }
    const double dv = canon_of(q0, n, 1, part), cv = canon_of(q1, n, 1, part);
    const double pobj = canon_of(q2, n, 0, part), dobjc = canon_of(q3, n, 0, part);
    const double dobj = dobjr + dobjc;
    const double ep = fmax(rv, cv) / (1.0 + bnorm);
    const double ed = fmax(dv, sv) / (1.0 + cnorm);
    const double eg = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
    out[0] = ep;
    out[1] = ed;
    out[2] = eg;
    out[3] = (isnan(ep) || isnan(ed) || isnan(eg)) ? ep + ed + eg : fmax(ep, fmax(ed, eg));
    out[4] = pobj;
    out[5] = dobj;
}

/* The CSR structure of the CSC input (columns ascending within a row), the
 * CSC -> CSR position map and the FP64 values in CSR order. */
static void build_csr(enh_lp *P, int m, int n, const int *col_ptr, const int *row_ind, const double *vals) {
    const int nnz = col_ptr[n];
    P->m = m; P->n = n; P->cp = col_ptr; P->ri = row_ind; P->v = vals;
    P->rp = (int *)calloc((size_t)m + 1, sizeof(int));
    P->ci = (int *)xmalloc(sizeof(int) * (size_t)nnz);
    P->c2r = (int *)xmalloc(sizeof(int) * (size_t)nnz);
    for (int k = 0; k < nnz; ++k) P->rp[row_ind[k] + 1]++;
    for (int i = 0; i < m; ++i) P->rp[i + 1] += P->rp[i];
    int *f = (int *)xmalloc(sizeof(int) * (size_t)(m + 1));
    memcpy(f, P->rp, sizeof(int) * (size_t)(m + 1));
    for (int j = 0; j < n; ++j)
        for (int k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
            const int d = f[row_ind[k]]++;
            P->ci[d] = j;
            P->c2r[k] = d;
        }
    free(f);
    P->rv = (double *)xmalloc(sizeof(double) * (size_t)nnz);
    for (int k = 0; k < nnz; ++k) P->rv[P->c2r[k]] = vals[k];
}

/* |b|_inf over the finite row bounds and |c|_inf, as the engine sets them up
 * (a NaN operand is skipped, as std::max(m, v) skips it when m is not less
 * than NaN). */
static void bc_norms(int m, int n, const double *c, const double *rl, const double *ru, double *bnorm,
                     double *cnorm) {
    double b = 0.0, cn = 0.0;
    for (int i = 0; i < m; ++i) {
        if (isfinite(rl[i]) && b < fabs(rl[i])) b = fabs(rl[i]);
        if (isfinite(ru[i]) && b < fabs(ru[i])) b = fabs(ru[i]);
    }
    for (int j = 0; j < n; ++j)
        if (cn < fabs(c[j])) cn = fabs(c[j]);
    *bnorm = b;
    *cnorm = cn;
}

double orc_canon_sum(const double *val, int n, int is_max) {
    double *part = (double *)xmalloc(sizeof(double) * (size_t)(n / ORC_GROUP + 1));
    const double r = canon_of(val, n, is_max, part);
    free(part);
    return r;
}

int orc_kkt_engine(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
                   const double *c, const double *l, const double *u, const double *rl,
                   const double *ru, const double *x, const double *y, double *out) {
    enh_lp P;
    build_csr(&P, m, n, col_ptr, row_ind, vals);
    double bnorm, cnorm;
    bc_norms(m, n, c, rl, ru, &bnorm, &cnorm);
    const int mx = (m > n ? m : n) + 1;
    double *w = (double *)xmalloc(sizeof(double) * 5 * (size_t)mx);
    double *ax = (double *)xmalloc(sizeof(double) * ((size_t)m + 1));
    double *lam = (double *)xmalloc(sizeof(double) * ((size_t)n + 1));
    kkt_full(&P, c, l, u, rl, ru, x, y, bnorm, cnorm, ax, lam, out, w);
    free(w); free(ax); free(lam);
    free(P.rp); free(P.ci); free(P.c2r); free(P.rv);
    return 0;
}

static double finite_or_zero(float a, float b) {
    return isfinite(a) ? (double)a : (isfinite(b) ? (double)b : 0.0);
}

int orc_pdhg_enh(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
                 const double *c, const double *l, const double *u, const double *rl,
                 const double *ru, const orc_enh_cfg *cfg, double *x_out, double *y_out,
                 orc_enh_out *out) {
    memset(out, 0, sizeof(*out));
    out->status = 8;
    if (m == 0 || n == 0) { out->status = 0; return 0; }
    enh_lp P;
    const int nnz = col_ptr[n];
    build_csr(&P, m, n, col_ptr, row_ind, vals);
    P.dr = (double *)xmalloc(sizeof(double) * (size_t)m);
    P.dc = (double *)xmalloc(sizeof(double) * (size_t)n);
    P.cv_s = (float *)xmalloc(sizeof(float) * (size_t)nnz);
    P.rv_s = (float *)xmalloc(sizeof(float) * (size_t)nnz);
    scale_lp(&P, cfg->ruiz_iters);

    /* round data (scaled, FP32): costs, bounds, dual offset */
    float *cs = (float *)xmalloc(sizeof(float) * (size_t)n), *ls = (float *)xmalloc(sizeof(float) * (size_t)n);
    float *us = (float *)xmalloc(sizeof(float) * (size_t)n), *rls = (float *)xmalloc(sizeof(float) * (size_t)m);
    float *rus = (float *)xmalloc(sizeof(float) * (size_t)m), *ybar = (float *)calloc((size_t)m, sizeof(float));
    for (int j = 0; j < n; ++j) {
        cs[j] = (float)(c[j] * P.dc[j]);
        ls[j] = (float)(l[j] / P.dc[j]);
        us[j] = (float)(u[j] / P.dc[j]);
    }
    for (int i = 0; i < m; ++i) {
        rls[i] = (float)(rl[i] * P.dr[i]);
        rus[i] = (float)(ru[i] * P.dr[i]);
    }
    /* FP64 base point of the refinement and its scales */
    double *xbase = (double *)calloc((size_t)n, sizeof(double));
    double *ybase = (double *)calloc((size_t)m, sizeof(double));
    double inv_ap = 1.0, inv_ad = 1.0, alpha_p = 1.0, alpha_d = 1.0;

    double bnorm, cnorm;
    bc_norms(m, n, c, rl, ru, &bnorm, &cnorm);
    const int mx = m > n ? m : n;
    double *kw = (double *)xmalloc(sizeof(double) * 5 * (size_t)mx);
    double *rq = (double *)xmalloc(sizeof(double) * 4 * (size_t)mx); /* per-element residual terms */
    double *rq0 = rq, *rq1 = rq + mx, *rq2 = rq + 2 * (size_t)mx, *rq3 = rq + 3 * (size_t)mx;
    const double sig = power_sigma(&P, 64);
    const double eta = (cfg->step > 0.0) ? cfg->step : 0.998 / (sig > 1e-12 ? sig : 1.0);
    double cn = 0.0, bn = 0.0;
    for (int j = 0; j < n; ++j) cn += (double)cs[j] * cs[j];
    for (int i = 0; i < m; ++i) { double b = finite_or_zero(rus[i], rls[i]); bn += b * b; }
    cn = sqrt(cn);
    bn = sqrt(bn);
    double omega = cfg->omega_start > 0.0 ? cfg->omega_start : ((cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0);

    float *x = (float *)xmalloc(sizeof(float) * (size_t)n), *x0 = (float *)xmalloc(sizeof(float) * (size_t)n);
    float *xp = (float *)xmalloc(sizeof(float) * (size_t)n), *xb = (float *)xmalloc(sizeof(float) * (size_t)n);
    float *y = (float *)calloc((size_t)m, sizeof(float)), *y0 = (float *)calloc((size_t)m, sizeof(float));
    float *yp = (float *)calloc((size_t)m, sizeof(float));
    float *g = (float *)xmalloc(sizeof(float) * (size_t)n), *h = (float *)xmalloc(sizeof(float) * (size_t)m);
    double *xc = (double *)xmalloc(sizeof(double) * (size_t)n), *yc_ = (double *)xmalloc(sizeof(double) * (size_t)m);
    double *ax = (double *)xmalloc(sizeof(double) * (size_t)m), *lam = (double *)xmalloc(sizeof(double) * (size_t)n);
    double *xbest = (double *)xmalloc(sizeof(double) * (size_t)n), *ybest = (double *)calloc((size_t)m, sizeof(double));
    double *axbest = (double *)xmalloc(sizeof(double) * (size_t)m), *lambest = (double *)xmalloc(sizeof(double) * (size_t)n);
    /* start point: the (warm) start in the scaled space, x clamped to the box */
    for (int j = 0; j < n; ++j) {
        float v = cfg->x_start ? (float)(cfg->x_start[j] / P.dc[j]) : 0.0f;
        if (v < ls[j]) v = ls[j];
        if (v > us[j]) v = us[j];
        x[j] = x0[j] = xp[j] = v;
    }
    if (cfg->y_start)
        for (int i = 0; i < m; ++i) y[i] = y0[i] = yp[i] = (float)(cfg->y_start[i] / P.dr[i]);
    const float rho = (float)cfg->reflection, rho1 = 1.0f + rho;
    const int F = cfg->check_frequency;
    int last_restart = 0, restarts = 0, restart_next = 0, rounds = 0;
    double pw_int = 0.0; /* primal-weight controller: sum of log-balance errors */
    double r_anchor = -1.0, r_prev = -1.0;
    double best[10] = {0}, kk[10] = {0};
    /* Refinement baseline: the first check's KKT (INFINITY: not yet taken).  A
     * warm start that carries its primal weight is a previous solve's point,
     * already at its FP32 iterates' precision: DBL_MAX keeps the first check
     * from setting the baseline, so the first restart re-centres the point in
     * FP64 (engine_enh.cpp sets the same per instance). */
    double best_kkt = INFINITY, round_start_kkt = cfg->omega_start > 0.0 ? DBL_MAX : INFINITY;
    int it, status = 4, have_best = 0;
    for (it = 0; it < cfg->max_iterations; ++it) {
        if (restart_next) { /* take z+ as iterate and anchor */
            memcpy(x, xp, sizeof(float) * (size_t)n);
            memcpy(x0, xp, sizeof(float) * (size_t)n);
            memcpy(y, yp, sizeof(float) * (size_t)m);
            memcpy(y0, yp, sizeof(float) * (size_t)m);
            restart_next = 0;
        }
        const float tau = (float)(eta / omega), sigma = (float)(eta * omega);
        const float inv_sigma = (float)(1.0 / (eta * omega));
        const int k = it - last_restart;
        const float a = (float)(k + 1) / (float)(k + 2), b = 1.0f / (float)(k + 2);
        const int check = ((it + 1) % F == 0) || (it + 1 == cfg->max_iterations);
        spmv32(n, P.cp, P.ri, P.cv_s, y, g);
        for (int j = 0; j < n; ++j) {
            const float xo = x[j];
            const float t = cs[j] + g[j];
            float v = fmaf(-tau, t, xo);
            if (v < ls[j]) v = ls[j];
            if (v > us[j]) v = us[j];
            xb[j] = 2.0f * v - xo;
            const float r = fmaf(rho1, v, -(rho * xo));
            x[j] = fmaf(a, r, b * x0[j]);
            if (check) {
                xp[j] = v;
                const double d = (double)xo - (double)v, e = (double)v - (double)x0[j];
                rq0[j] = d * d;
                rq1[j] = e * e;
            }
        }
        spmv32(m, P.rp, P.ci, P.rv_s, xb, h);
        for (int i = 0; i < m; ++i) {
            /* y+ = sigma (w - clamp(w, rl, ru)) - ybar with w = h + (y + ybar) / sigma,
             * evaluated as y + sigma (h - bound) on an active bound (exact
             * subtraction near the bound) and -ybar in between. */
            const float yo = y[i];
            const float w = fmaf(yo + ybar[i], inv_sigma, h[i]);
            float yn = -ybar[i];
            if (w < rls[i]) yn = fmaf(sigma, h[i] - rls[i], yo);
            else if (w > rus[i]) yn = fmaf(sigma, h[i] - rus[i], yo);
            const float r = fmaf(rho1, yn, -(rho * yo));
            y[i] = fmaf(a, r, b * y0[i]);
            if (check) {
                yp[i] = yn;
                const double d = (double)yo - (double)yn, e = (double)yn - (double)y0[i];
                rq2[i] = d * d;
                rq3[i] = e * e;
            }
        }
        if (!check) continue;
        double *part = kw; /* the KKT below reuses the scratch */
        const double dx2 = canon_of(rq0, n, 0, part), dxa2 = canon_of(rq1, n, 0, part);
        const double dy2 = canon_of(rq2, m, 0, part), dya2 = canon_of(rq3, m, 0, part);
        /* ---- check: FP64 KKT of the unscaled T-output ---- */
        for (int j = 0; j < n; ++j) xc[j] = xbase[j] + P.dc[j] * (double)xp[j] * inv_ap;
        for (int i = 0; i < m; ++i) yc_[i] = ybase[i] + P.dr[i] * (double)yp[i] * inv_ad;
        kkt_full(&P, c, l, u, rl, ru, xc, yc_, bnorm, cnorm, ax, lam, kk, kw);
        const int done = it + 1;
        if (isnan(kk[3])) { status = 6; it = done; break; } /* NumericalFailure */
        if (!(kk[3] >= best_kkt)) { /* improvement (NaN never counts) */
            best_kkt = kk[3];
            memcpy(best, kk, sizeof(kk));
            memcpy(xbest, xc, sizeof(double) * (size_t)n);
            memcpy(ybest, yc_, sizeof(double) * (size_t)m);
            memcpy(axbest, ax, sizeof(double) * (size_t)m);
            memcpy(lambest, lam, sizeof(double) * (size_t)n);
            have_best = 1;
        }
        if (kk[3] <= cfg->tol) { status = 0; it = done; break; }
        if (done >= cfg->max_iterations) { it = done; break; }
        if (round_start_kkt == INFINITY) round_start_kkt = kk[3]; /* first check of the solve */
        const double r = sqrt(omega * dx2 + dy2 / omega);
        const int kc = done - last_restart;
        if (r_anchor < 0.0) { r_anchor = r_prev = r; continue; }
        const int rs = (r <= 0.2 * r_anchor) || (r <= 0.8 * r_anchor && r > r_prev) ||
                       ((double)kc >= 0.36 * (double)done);
        r_prev = r;
        if (!rs) continue;
        /* ---- restart to z+ = this check's point ---- */
        {
            const double ddx = sqrt(dxa2), ddy = sqrt(dya2);
            if (ddx > 1e-10 && ddy > 1e-10) {
                const double e = log(omega * ddx / ddy);
                pw_int += e;
                omega = exp(log(omega) - (0.6 * e + 0.05 * pw_int));
            }
        }
        last_restart = done;
        restarts++;
        r_anchor = r;
        if (rounds < cfg->max_refinements && kk[3] <= cfg->refine_ratio * round_start_kkt) {
            /* Refinement at the restart: the same restart point, re-centred in
             * FP64 as the residual LP (x = x~ + D_c d / alpha_p,
             * y = y~ + D_r e / alpha_d) so that the FP32 iterates carry the
             * correction at full precision.  In exact arithmetic this is the
             * plain restart (same effective primal weight). */
            const double dp = fmax(kk[0], kk[2]), dd = fmax(kk[1], kk[2]);
            double ap = fmin(1.0 / fmax(dp, 1e-300), 1e12), ad = fmin(1.0 / fmax(dd, 1e-300), 1e12);
            if (ap < 1.0) ap = 1.0;
            if (ad < 1.0) ad = 1.0;
            omega *= (ad / ap) / (alpha_d / alpha_p);
            alpha_p = ap; alpha_d = ad; inv_ap = 1.0 / ap; inv_ad = 1.0 / ad;
            memcpy(xbase, xc, sizeof(double) * (size_t)n);
            memcpy(ybase, yc_, sizeof(double) * (size_t)m);
            for (int j = 0; j < n; ++j) {
                cs[j] = (float)(P.dc[j] * (ad * lam[j]));
                ls[j] = (float)(ap * (l[j] - xbase[j]) / P.dc[j]);
                us[j] = (float)(ap * (u[j] - xbase[j]) / P.dc[j]);
                float v = 0.0f;
                if (v < ls[j]) v = ls[j];
                if (v > us[j]) v = us[j];
                x[j] = x0[j] = xp[j] = v;
            }
            for (int i = 0; i < m; ++i) {
                rls[i] = (float)(ap * (rl[i] - ax[i]) * P.dr[i]);
                rus[i] = (float)(ap * (ru[i] - ax[i]) * P.dr[i]);
                ybar[i] = (float)(ad * ybase[i] / P.dr[i]);
                y[i] = y0[i] = yp[i] = 0.0f;
            }
            rounds++;
            round_start_kkt = kk[3];
            continue;
        }
        restart_next = 1;
    }
    if (status != 0 && !have_best) {
        for (int j = 0; j < n; ++j) xbest[j] = xbase[j] + P.dc[j] * (double)xp[j] * inv_ap;
        for (int i = 0; i < m; ++i) ybest[i] = ybase[i] + P.dr[i] * (double)yp[i] * inv_ad;
        kkt_full(&P, c, l, u, rl, ru, xbest, ybest, bnorm, cnorm, axbest, lambest, best, kw);
    }
    out->status = status;
    out->iterations = it;
    out->restarts = restarts;
    out->refinements = rounds;
    out->eps_p = best[0]; out->eps_d = best[1]; out->eps_gap = best[2]; out->kkt = best[3];
    out->primal_obj = best[4]; out->dual_obj = best[5];
    out->step = eta;
    out->omega = omega * alpha_p / alpha_d; /* effective weight (alpha = 1 frame) */
    if (x_out) memcpy(x_out, xbest, sizeof(double) * (size_t)n);
    if (y_out) memcpy(y_out, ybest, sizeof(double) * (size_t)m);
    free(kw); free(rq);
    free(P.rp); free(P.ci); free(P.c2r); free(P.rv); free(P.dr); free(P.dc); free(P.cv_s); free(P.rv_s);
    free(cs); free(ls); free(us); free(rls); free(rus); free(ybar); free(xbase); free(ybase);
    free(x); free(x0); free(xp); free(xb); free(y); free(y0); free(yp); free(g); free(h);
    free(xc); free(yc_); free(ax); free(lam); free(xbest); free(ybest); free(axbest); free(lambest);
    return 0;
}
