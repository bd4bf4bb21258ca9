/*
 * oracle/pdhg_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * The oracle: a plain-C CPU restatement of the reference GPU PDHG path of
 * refinery-optimizer (core/src/gpu/gpu_solver.cpp:60-417,
 * core/src/gpu/cuda_pdhg.cu:14-151) plus the FP64 relative-KKT evaluator
 * and a CPU port of this engine's enhanced (restarted Halpern) mode.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and only as the checker or the timed CPU baseline.  Nothing in the
 * product (refinery-optimizer_amd/) includes, links or calls it.
 *
 * Pinning (what keeps each function honest; see the tests named):
 *   orc_csc_to_csr          dense brute force + CSR^T==CSC invariants
 *                            (tests/test_oracle.py)
 *   orc_spmv_*              3x3 KATs of tests/unit/test_sparse_matrix.cpp:45-63,
 *                            dense numpy brute force
 *   orc_ref_setup_*         closed forms (0.95/max(1,|A|_1)), numpy recompute
 *   orc_pdhg_ref_*          toy01 closed form (x*=(1.5,2.5), obj -6.5,
 *                            data/toy/toy_suite_metadata.json:5-12), reference
 *                            C++ dual-simplex objectives (oracle/_ref),
 *                            scipy HiGHS objectives, first-iterate closed form,
 *                            restart-kernel no-op invariant, Appendix-B KAT of
 *                            the reference CSC upload defect
 *   orc_kkt                 hand-built optimal / perturbed points, HiGHS duals
 *   orc_identify_basis      hand-computed rankings on tiny inputs
 *   orc_pdhg_enh            FP64 KKT of its output + HiGHS objectives; its
 *                            row sums by the split-rule KATs (long rows)
 *   orc_canon_sum           sums whose results rounding alone fixes (2^53
 *                            plus ones the lane butterfly pairs first, +-2^53
 *                            in one lane, waves added in order), exact integer
 *                            sums (tests/test_oracle.py)
 *   orc_kkt_engine          toy01 closed forms, orc_kkt to 1e-12
 */
#ifndef PDHG_ORACLE_H
#define PDHG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ref_cfg {
    int max_iterations;
    double tol_primal, tol_dual;
    int check_frequency;
    double step_tau, step_sigma;
    int enable_restart;
    /* Optional start point (an extension: the reference always starts from
     * clamp(0, l, u) and y = 0): x0[n] cast to R and clamped to the bounds in
     * place of 0, y0[m] (PDHG sign) cast to R in place of 0.  NULL: cold. */
    const double *x0, *y0;
} orc_ref_cfg;

typedef struct orc_ref_out {
    int status;       /* model::SolutionStatus values: 0 Optimal, 4 IterationLimit */
    int iterations;   /* the reference's loop counter at exit (K+1 when not converged) */
    int checks;
    double last_primal, last_dual; /* reference P and D at the last check */
    double best_residual;
    double objective;  /* sum_j c_j x_best_j in FP64 */
    double max_abs_c, max_abs_b;
} orc_ref_out;

int orc_csc_to_csr(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
                   int *row_ptr, int *col_ind, double *csr_vals);

void orc_spmv_csr_f64(int m, const int *row_ptr, const int *col_ind, const double *vals,
                      const double *x, double *y);
void orc_spmv_csc_f64(int n, const int *col_ptr, const int *row_ind, const double *vals,
                      const double *y, double *z);
void orc_spmv_seq_f32(int rows, const int *ptr, const int *idx, const float *vals,
                      const float *x, float *out);
/* The enhanced mode's FP32 row sums (pdhg_enh.c): sequential chains, with
 * rows of more than 8192 nonzeros summed in 64-nonzero segments within
 * 4096-nonzero blocks. */
void orc_spmv_enh_f32(int rows, const int *ptr, const int *idx, const float *vals,
                      const float *x, float *out);

void orc_ref_setup_f32(int m, int n, const int *col_ptr, const double *vals_d,
                       const int *row_ptr, const float *csr_vals, const double *c,
                       const double *rl, const double *ru, float step_tau, float step_sigma,
                       float *tau, float *sigma, float *max_abs_c, float *max_abs_b);

int orc_pdhg_ref_f32(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
                     const double *c, const double *l, const double *u, const double *rl,
                     const double *ru, const orc_ref_cfg *cfg, double *x_out, double *y_out,
                     float *x_last, float *xbar_last, orc_ref_out *out);
int orc_pdhg_ref_f64(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
                     const double *c, const double *l, const double *u, const double *rl,
                     const double *ru, const orc_ref_cfg *cfg, double *x_out, double *y_out,
                     double *x_last, double *xbar_last, orc_ref_out *out);

/* Relative KKT (FP64, l-infinity), reference dual sign: lambda = c + A^T y.
 * out[0..9] = eps_p, eps_d, eps_gap, kkt, primal_obj, dual_obj,
 *             max_row_viol, max_col_viol, max_dual_viol, |b|_inf */
int orc_kkt(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
            const double *c, const double *l, const double *u, const double *rl,
            const double *ru, const double *x, const double *y, double *out);

/* gpu_solver.cpp:319-395 basis identification (stable order on ties). */
int orc_identify_basis(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
                       const double *l, const double *u, const double *rl, const double *ru,
                       const double *x, int8_t *col_basis, int8_t *row_basis);

/* The enhanced mode's fixed-order FP64 total of val[0, n): 16-element group
 * partials by a lane butterfly, then the 256-lane instance tree (sum, or max
 * with is_max). */
double orc_canon_sum(const double *val, int n, int is_max);

/* The enhanced mode's FP64 check of (x, y) on the original LP, in the
 * engine's order: out[0..5] = eps_p, eps_d, eps_gap, kkt, primal_obj,
 * dual_obj (|b|_inf over finite row bounds, |c|_inf). */
int orc_kkt_engine(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
                   const double *c, const double *l, const double *u, const double *rl,
                   const double *ru, const double *x, const double *y, double *out);

/* CPU port of the enhanced mode (Ruiz + Pock-Chambolle scaling, power-iteration
 * step, reflected restarted Halpern PDHG, FP64 checks).  No reference
 * counterpart.  Returns status 0 (converged to tol) or 4. */
typedef struct orc_enh_cfg {
    int max_iterations;
    int check_frequency;
    double tol;            /* relative KKT target */
    double reflection;     /* rho in [0,1] */
    int ruiz_iters;
    int threads;           /* OpenMP threads (1 = scalar) */
    double step;           /* eta; <= 0: 0.998 / sigma_max by power iteration */
    int max_refinements;   /* outer FP64 refinement rounds */
    double refine_ratio;   /* at a restart, refine once KKT <= ratio * KKT at the last re-centring */
    const double *x_start; /* [n] warm start (engine: mpdhg_warm_start.x), or NULL: clamp(0, l, u) */
    const double *y_start; /* [m] warm start in the reference dual sign, or NULL: 0 */
    double omega_start;    /* primal weight to start from; <= 0: |c| / |b| of the scaled LP */
} orc_enh_cfg;
typedef struct orc_enh_out {
    int status, iterations, restarts, refinements;
    double eps_p, eps_d, eps_gap, kkt, primal_obj, dual_obj, step;
    double omega;      /* the effective primal weight at the end, omega alpha_p / alpha_d
                        * (the frame a new solve starts in: a warm start's omega_start) */
} orc_enh_out;
int orc_pdhg_enh(int m, int n, const int *col_ptr, const int *row_ind, const double *vals,
                 const double *c, const double *l, const double *u, const double *rl,
                 const double *ru, const orc_enh_cfg *cfg, double *x_out, double *y_out,
                 orc_enh_out *out);

#ifdef __cplusplus
}
#endif
#endif
