/*
 * oracle/pdhg_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of the reference GPU PDHG path (refinery-optimizer,
 * core/src/gpu/gpu_solver.cpp and core/src/gpu/cuda_pdhg.cu) and of the
 * residual semantics it is judged by.  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg only.  Build:
 * oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include "pdhg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- reference-mode loop, FP32 mirror and FP64 variant ---------------- */
#define R float
#define RFMA fmaf
#define RABS fabsf
#define RMAXF fmaxf
#define RSFX f32
#define R_BIG 1e15f
#define R_TAUMIN 1e-4f
#include "pdhg_ref_body.h"
#undef R
#undef RFMA
#undef RABS
#undef RMAXF
#undef RSFX
#undef R_BIG
#undef R_TAUMIN

#define R double
#define RFMA fma
#define RABS fabs
#define RMAXF fmax
#define RSFX f64
#define R_BIG 1e15
#define R_TAUMIN 1e-4
#include "pdhg_ref_body.h"
#undef R
#undef RFMA
#undef RABS
#undef RMAXF
#undef RSFX
#undef R_BIG
#undef R_TAUMIN

void orc_spmv_seq_f32(int rows, const int *ptr, const int *idx, const float *vals,
                      const float *x, float *out) {
    seq_spmv_f32(rows, ptr, idx, vals, x, out);
}

/* gpu_solver.cpp:92-109: counting sort CSC -> CSR, columns ascending per row. */
int orc_csc_to_csr(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
                   int *row_ptr, int *col_ind, double *csr_vals) {
    const int nnz = col_ptr[n];
    memset(row_ptr, 0, sizeof(int) * (size_t)(m + 1));
    for (int k = 0; k < nnz; ++k) {
        if (row_ind[k] < 0 || row_ind[k] >= m) return -1;
        row_ptr[row_ind[k] + 1]++;
    }
    for (int i = 0; i < m; ++i) row_ptr[i + 1] += row_ptr[i];
    int *fill = (int *)malloc(sizeof(int) * (size_t)(m + 1));
    memcpy(fill, row_ptr, sizeof(int) * (size_t)(m + 1));
    for (int j = 0; j < n; ++j)
        for (int k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
            int d = fill[row_ind[k]]++;
            col_ind[d] = j;
            csr_vals[d] = vals[k];
        }
    free(fill);
    return 0;
}

/* core/src/model/sparse_matrix.cpp:90-104 (mat_vec) and :106-129 (mat_trans_vec). */
void orc_spmv_csr_f64(int m, const int *row_ptr, const int *col_ind, const double *vals,
                      const double *x, double *y) {
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) s += vals[k] * x[col_ind[k]];
        y[i] = s;
    }
}
void orc_spmv_csc_f64(int n, const int *col_ptr, const int *row_ind, const double *vals,
                      const double *y, double *z) {
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int k = col_ptr[j]; k < col_ptr[j + 1]; ++k) s += vals[k] * y[row_ind[k]];
        z[j] = s;
    }
}

/* ---- FP64 relative KKT ------------------------------------------------
 * Generalises docs/gpu_pdhg.md:49-57 to row ranges and boxes, in the
 * reference PDHG's dual sign (lambda = c + A^T y, the sign of
 * cuda_pdhg.cu:65 and :82-103).  Dual objective:
 *   D = sum_j (l_j lam_j^+ - u_j lam_j^-) - sum_i (ru_i y_i^+ - rl_i y_i^-)
 * evaluated on the projection of lambda onto the sign set the bounds allow. */
static inline int fin(double v) { return isfinite(v); }

/* Row duals carry signs too: y_i > 0 needs a finite ru_i, y_i < 0 a finite
 * rl_i (the dual prox of cuda_pdhg.cu:96-101 always lands there).  A point
 * that breaks this is dual infeasible by |y_i - proj(y_i)|, and the dual
 * objective is evaluated at the projection. */
static inline double proj_row_dual(double y, double rl, double ru) {
    if (!fin(ru) && y > 0.0) return 0.0;
    if (!fin(rl) && y < 0.0) return 0.0;
    return y;
}

double orc_dual_row_term(double y, double rl, double ru) {
    y = proj_row_dual(y, rl, ru);
    if (y > 0.0) return -ru * y;
    if (y < 0.0) return -rl * y;
    return 0.0;
}

int orc_kkt(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
            const double *c, const double *l, const double *u, const double *rl,
            const double *ru, const double *x, const double *y, double *out) {
    double *ax = (double *)calloc((size_t)(m ? m : 1), sizeof(double));
    /* column-major scatter A x */
    for (int j = 0; j < n; ++j)
        for (int k = col_ptr[j]; k < col_ptr[j + 1]; ++k) ax[row_ind[k]] += vals[k] * x[j];
    double bn = 0.0, cn = 0.0, rowv = 0.0, colv = 0.0, dv = 0.0, pobj = 0.0, dobj = 0.0;
    for (int i = 0; i < m; ++i) {
        if (fin(rl[i])) bn = fmax(bn, fabs(rl[i]));
        if (fin(ru[i])) bn = fmax(bn, fabs(ru[i]));
        double v = 0.0;
        if (fin(rl[i]) && ax[i] < rl[i]) v = rl[i] - ax[i];
        if (fin(ru[i]) && ax[i] > ru[i]) v = ax[i] - ru[i];
        rowv = fmax(rowv, v);
        dv = fmax(dv, fabs(y[i] - proj_row_dual(y[i], rl[i], ru[i])));
        dobj += orc_dual_row_term(y[i], rl[i], ru[i]);
    }
    for (int j = 0; j < n; ++j) {
        cn = fmax(cn, fabs(c[j]));
        double v = 0.0;
        if (fin(l[j]) && x[j] < l[j]) v = l[j] - x[j];
        if (fin(u[j]) && x[j] > u[j]) v = x[j] - u[j];
        colv = fmax(colv, v);
        double lam = c[j];
        for (int k = col_ptr[j]; k < col_ptr[j + 1]; ++k) lam += vals[k] * y[row_ind[k]];
        double p = lam; /* projection onto the allowed sign set */
        if (!fin(l[j]) && p > 0.0) p = 0.0;
        if (!fin(u[j]) && p < 0.0) p = 0.0;
        dv = fmax(dv, fabs(lam - p));
        if (p > 0.0) dobj += l[j] * p;
        else if (p < 0.0) dobj += u[j] * p;
        pobj += c[j] * x[j];
    }
    double ep = fmax(rowv, colv) / (1.0 + bn);
    double ed = dv / (1.0 + cn);
    double eg = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
    out[0] = ep; out[1] = ed; out[2] = eg;
    out[3] = (isnan(ep) || isnan(ed) || isnan(eg)) ? ep + ed + eg : fmax(ep, fmax(ed, eg));
    out[4] = pobj; out[5] = dobj; out[6] = rowv; out[7] = colv; out[8] = dv; out[9] = bn;
    free(ax);
    return 0;
}

/* ---- gpu_solver.cpp:319-395: hybrid basis identification ----------- */
typedef struct { int id; double dist; int8_t nb; } orc_cand;

static int cand_cmp(const void *pa, const void *pb) {
    const orc_cand *a = (const orc_cand *)pa, *b = (const orc_cand *)pb;
    if (a->dist > b->dist) return -1;  /* descending distance (:371-373) */
    if (a->dist < b->dist) return 1;
    return (a->id < b->id) ? -1 : (a->id > b->id); /* stable: index tie-break */
}

int orc_identify_basis(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
                       const double *l, const double *u, const double *rl, const double *ru,
                       const double *x, int8_t *col_basis, int8_t *row_basis) {
    double *ax = (double *)calloc((size_t)(m ? m : 1), sizeof(double));
    for (int j = 0; j < n; ++j) { /* :320-330 (skip |x|<=1e-12) */
        double xc = x[j];
        if (fabs(xc) > 1e-12)
            for (int k = col_ptr[j]; k < col_ptr[j + 1]; ++k) ax[row_ind[k]] += vals[k] * xc;
    }
    orc_cand *cs = (orc_cand *)malloc(sizeof(orc_cand) * (size_t)(n + m + 1));
    for (int t = 0; t < n + m; ++t) { /* :342-368 */
        double val, lo, hi;
        if (t < n) { val = x[t]; lo = l[t]; hi = u[t]; }
        else { val = ax[t - n]; lo = rl[t - n]; hi = ru[t - n]; }
        double dl = (lo > -1e15) ? fmax(0.0, val - lo) : 1e30;
        double du = (hi < 1e15) ? fmax(0.0, hi - val) : 1e30;
        cs[t].id = t;
        cs[t].dist = fmin(dl, du);
        cs[t].nb = (dl <= du) ? 1 : 2; /* AtLower : AtUpper */
    }
    qsort(cs, (size_t)(n + m), sizeof(orc_cand), cand_cmp);
    for (int k = 0; k < n + m; ++k) { /* :375-395 */
        int id = cs[k].id;
        int8_t st = (k < m) ? 0 /* Basic */ : cs[k].nb;
        if (id < n) col_basis[id] = st; else row_basis[id - n] = st;
    }
    free(cs);
    free(ax);
    return 0;
}
