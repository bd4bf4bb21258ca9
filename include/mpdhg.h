/*
 * mpdhg.h -- C ABI of the MI355X PDHG LP engine (libmpdhg.so).
 *
 * Drop-in boundary for the GPU PDHG path of refinery-optimizer:
 *   sih::gpu::GpuSolver            core/include/sih/gpu/gpu_solver.hpp:30-45
 *   sih::gpu::GpuPdhgConfig        core/include/sih/gpu/gpu_solver.hpp:12-22
 *   sih::gpu::GpuMemoryInfo        core/include/sih/gpu/gpu_solver.hpp:24-28
 *   sih::model::Solution fields    core/include/sih/model/solution.hpp:37-84
 *   sih::model::SparseMatrix CSC   core/include/sih/model/sparse_matrix.hpp:38-40
 * Plain C types only: int64 CSC indices and double values are taken exactly
 * as model::SparseMatrix stores them, so a caller passes
 * A().csc_col_ptr().data() etc. without conversion.
 *
 * Errors: every entry point returns 0 on success and a negative MPDHG_E_*
 * code on failure; mpdhg_last_error() gives a thread-local message.  No C++
 * exception crosses this boundary.
 * Threading: a context owns one HIP stream and its device buffers; use one
 * context per host thread.  Calls on different contexts may run concurrently.
 */
#ifndef MPDHG_H
#define MPDHG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define MPDHG_OK 0
#define MPDHG_E_ARG (-1)      /* invalid argument / malformed LP */
#define MPDHG_E_HIP (-2)      /* HIP runtime error (message has the call) */
#define MPDHG_E_NODEV (-3)    /* no HIP device visible */
#define MPDHG_E_NOMEM (-4)    /* device or host allocation failed */
#define MPDHG_E_STATE (-5)    /* call not valid in the context's state */

/* ---- model::SolutionStatus values (solution.hpp:11-21) ---------------- */
#define MPDHG_STATUS_OPTIMAL 0
#define MPDHG_STATUS_INFEASIBLE 1
#define MPDHG_STATUS_UNBOUNDED 2
#define MPDHG_STATUS_TIME_LIMIT 3
#define MPDHG_STATUS_ITERATION_LIMIT 4
#define MPDHG_STATUS_NUMERICAL_FAILURE 6
#define MPDHG_STATUS_UNKNOWN 8

/* ---- model::BasisStatus values (solution.hpp:23-29) ------------------- */
#define MPDHG_BASIS_BASIC 0
#define MPDHG_BASIS_AT_LOWER 1
#define MPDHG_BASIS_AT_UPPER 2

/* ---- solver modes (extension) ----------------------------------------- */
#define MPDHG_MODE_REFERENCE 0 /* the reference algorithm, FP32, bit-exact restatement */
#define MPDHG_MODE_ENHANCED 1  /* scaled, restarted reflected-Halpern PDHG + FP64 checks */

/* GpuPdhgConfig (gpu_solver.hpp:12-22): the first nine fields keep the
 * reference names, meanings and defaults; the rest are extensions. */
typedef struct mpdhg_config {
    int32_t max_iterations;       /* 4000 */
    float tol_primal;             /* 1e-4 */
    float tol_dual;               /* 1e-4 */
    float tol_gap;                /* 1e-4 (reference: unread; enhanced: gap part of KKT) */
    int32_t check_frequency;      /* 25 */
    float step_tau;               /* 0 = automatic */
    float step_sigma;             /* 0 = automatic */
    int32_t enable_restart;       /* 1 (reference mode: provably without effect) */
    int32_t enable_hybrid_polish; /* 1 (reference: unread) */
    /* extensions */
    int32_t mode;                 /* MPDHG_MODE_REFERENCE */
    double tol_kkt;               /* enhanced: FP64 relative-KKT target; <=0: max(tol_primal,tol_dual,tol_gap) */
    int32_t max_refinements;      /* enhanced: outer FP64 iterative-refinement rounds (0 = off) */
    float reflection;             /* enhanced: Halpern reflection coefficient rho in [0,1] (1.0) */
    int32_t ruiz_iterations;      /* enhanced: Ruiz equilibration passes (10) */
    int32_t profile;              /* 1: time the fused SpMV kernels with HIP events */
    int32_t dual_sign;            /* 0: reference PDHG sign (reduced cost = c + A^T y);
                                     1: simplex/checker sign (c - A^T y, checker.cpp:231) */
    double refine_ratio;          /* enhanced: at a restart, re-centre the FP32 iterates on the
                                     restart point (one refinement round) once the FP64 KKT fell
                                     below ratio * its value at the last re-centring (1e-2) */
    int32_t spmv_layout;          /* 0: auto -- each LDS chunk of nonzeros stored sorted by
                                     gathered index where every instance's vectors are short
                                     enough for the packed index word (< 2^32 / chunk entries);
                                     1: storage order everywhere. */
    int32_t spmv_chunk;           /* nonzeros staged through LDS per workgroup pass: 0 = auto,
                                     or 1024, 2048, 4096, 8192 */
    int32_t launch_graph;         /* 1 = record one check cycle (the fused launch pairs,
                                     the FP64 check, the decisions) as a HIP graph and replay it
                                     for every full cycle; 0 = launch kernel by kernel. */
    int32_t spmv_block;           /* threads per SpMV workgroup (rows per tile at most): 0 = auto,
                                     or 128, 256, 512.  (spmv_block, spmv_chunk) must be one of
                                     (128,1024) (128,2048) (256,1024) (256,2048) (256,4096)
                                     (512,4096) (512,8192) when both are set.  Layout, chunk,
                                     block and launch_graph change speed only: every result is
                                     bit-identical for every choice (rows are summed in storage
                                     order, and every FP64 sum behind a decision is reduced per
                                     16-row group and then per instance in a fixed order). */
    int32_t spmv_tune;            /* enhanced and reference modes: 1 (default) = time every
                                     compiled schedule allowed by spmv_block / spmv_chunk on this
                                     batch's own matrix, per direction, and keep the fastest,
                                     when the batch holds >= 262144 nonzeros (smaller batches use
                                     the rule); 2 = always; 0 = never (the rule: the largest
                                     workgroup that still gives >= 1024 tiles). */
    int32_t cpu_fallback;         /* 0 (default): with no HIP device visible, mpdhg_solve /
                                     mpdhg_solve_batch / mpdhg_solve_hybrid return
                                     MPDHG_E_NODEV.  1: they do what GpuSolver::solve_pdhg does
                                     without a device (gpu_solver.cpp:66-71): print a notice to
                                     stderr and solve each LP with the host FP64 simplex
                                     (mpdhg_polish from the slack basis); status, iterations
                                     (simplex pivots), x, duals and objective then come from the
                                     simplex, the duals converted to this result's dual_sign. */
    int32_t vertex_check;         /* 0 (default): off.  k > 0: after the solve, every LP with at
                                     most k rows is taken from its PDHG point to an FP64 optimal
                                     vertex by the host crossover (mpdhg_identify_basis, then
                                     mpdhg_polish from that basis and point; 60 s per LP at most),
                                     and its result carries the vertex's status and objective and
                                     the PDHG objective's relative error against it. */
    int32_t spmv_bands;           /* one LP (either mode): split the A xbar product (CSR; in
                                     enhanced mode the fused dual kernel's) into
                                     this many column bands, one launch each, so that each band's
                                     slice of xbar stays in an XCD's 4 MB L2 (rows carry their
                                     running sums from band to band in storage order: results do
                                     not change).  0 (default) = by rule (bands of at most 3 MiB
                                     of xbar once xbar exceeds 3 MiB), 1 = no bands, k = k bands
                                     (at most 16). */
} mpdhg_config;

/* A model::Problem's LP data (problem.hpp:37-137) in CSC form.  Infinite
 * bounds are +-INFINITY (SIH_INFINITY).  The reference solve_pdhg ignores
 * sense() and obj_offset() (gpu_solver.cpp:300-302); so does this engine. */
typedef struct mpdhg_lp {
    int64_t m, n, nnz;
    const int64_t *col_ptr;  /* [n+1] */
    const int64_t *row_ind;  /* [nnz], rows ascending within a column */
    const double *vals;      /* [nnz] */
    const double *c;         /* [n] */
    const double *col_lower; /* [n] */
    const double *col_upper; /* [n] */
    const double *row_lower; /* [m] */
    const double *row_upper; /* [m] */
} mpdhg_lp;

/* model::Solution subset the GPU path fills (solution.hpp:37-84) plus
 * per-solve metrics.  Vector outputs are caller-allocated; NULL skips one. */
typedef struct mpdhg_result {
    double *x;             /* [n] primal (reference mode: best x, gpu_solver.cpp:261-264) */
    double *y;             /* [m] row duals (sign per mpdhg_config.dual_sign) */
    double *reduced_costs; /* [n] */
    double *row_activity;  /* [m] A x ("slack" in model::Solution) */
    int32_t status;        /* MPDHG_STATUS_* */
    int32_t restarts;
    int32_t refinements;
    int32_t checks;
    int64_t iterations;    /* reference: loop counter at exit (max_iterations+1 if not converged) */
    double primal_objective; /* sum_j c_j x_j (FP64) */
    double dual_objective;
    double rel_primal, rel_dual, rel_gap, rel_kkt; /* FP64 relative KKT of (x, y) */
    double ref_primal, ref_dual; /* reference-mode P and D at the last check */
    double time_wall_sec;    /* whole call, host setup and copies included */
    double solve_sec;        /* device-resident iteration loop only */
    double kernel_ms_spmv_csr, kernel_ms_spmv_csc; /* profile=1: summed event times */
    int64_t kernel_launches_csr, kernel_launches_csc;
    double bytes_csr_launch, bytes_csc_launch;     /* algorithmic bytes per fused launch */
    double step_size;        /* enhanced: eta (tau = eta/omega, sigma = eta*omega) */
    double primal_weight;    /* enhanced: final effective omega (the weight with the refinement
                              * rescaling undone; a warm start's primal_weight) */
    int32_t csr_block, csr_chunk;  /* tile schedule the solve used for A x (CSR) ... */
    int32_t csc_block, csc_chunk;  /* ... and for A^T y (CSC) */
    double tune_ms;                /* host + device time of the schedule search (0: rule) */
    int32_t csr_bands;             /* column bands of the A xbar product (1: none) */
    int32_t vertex_status;         /* vertex_check: MPDHG_STATUS_* of the crossover (UNKNOWN: not run) */
    double vertex_objective;       /* vertex_check: c^T x at the FP64 vertex (NaN: not run) */
    double vertex_rel_error;       /* vertex_check: |primal_objective - vertex_objective| /
                                      max(1, |vertex_objective|) (NaN unless the vertex is optimal) */
} mpdhg_result;

typedef struct mpdhg_ctx mpdhg_ctx;

/* ---- device queries: GpuSolver::is_cuda_available / get_vram_info ----- */
/* gpu_solver.cpp:42-46 */
int mpdhg_device_count(int *count);
/* gpu_solver.cpp:48-58; used = total - free */
int mpdhg_mem_info(int device, size_t *total_bytes, size_t *free_bytes, size_t *used_bytes);
const char *mpdhg_last_error(void);
void mpdhg_config_default(mpdhg_config *cfg);

/* ---- contexts (the reference allocates and frees per call,
 *      gpu_solver.cpp:153-216 and :276-288; a context keeps them) ------- */
int mpdhg_create(int device, mpdhg_ctx **out);
void mpdhg_destroy(mpdhg_ctx *ctx);

/* GpuSolver::solve_pdhg (gpu_solver.cpp:60-307).  ctx == NULL makes a
 * temporary context on device 0 (the reference's per-call lifetime). */
int mpdhg_solve(mpdhg_ctx *ctx, const mpdhg_lp *lp, const mpdhg_config *cfg,
                mpdhg_result *res);

/* No-allocation variant of mpdhg_solve.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory -- code running under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  Every
 * device buffer of the solve is carved from a scratch array of
 * MPDHG_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends; the call makes no hipMalloc, hipFree, pinned-host
 * allocation, event or graph (host heap memory is used for the host-side
 * CSR build).  ctx must come from mpdhg_create (which sets the scratch up on
 * its device).  The solve is mpdhg_solve's, launched kernel by kernel, with
 * the tile schedules chosen by rule and no column bands (spmv_tune,
 * spmv_bands and launch_graph are ignored, profile is off), which changes no
 * result: every result equals
 * mpdhg_solve's bit for bit.  An LP whose buffers could exceed the scratch is
 * rejected with MPDHG_E_ARG before any device work; reference mode needs about
 * 32 B per nonzero and 60 B per row and column, enhanced mode about 68 B per
 * nonzero and 116 B per row and column (so LPs up to ~1.5M and ~0.65M nonzeros
 * fit).  Calls on one GPU share its scratch and leave their buffers in it:
 * calls that overlap on one GPU (from several threads or contexts) race and
 * must be serialised by the caller. */
#define MPDHG_NOALLOC_SCRATCH_BYTES (64u << 20)
int mpdhg_solve_noalloc(mpdhg_ctx *ctx, const mpdhg_lp *lp, const mpdhg_config *cfg,
                        mpdhg_result *res);

/* Batched solve (extension): `count` LPs of any sizes in one device-resident
 * sweep; every launch covers all still-running instances.  Results are
 * per instance, identical to mpdhg_solve on each LP in reference mode. */
int mpdhg_solve_batch(mpdhg_ctx *ctx, const mpdhg_lp *lps, int count,
                      const mpdhg_config *cfg, mpdhg_result *res);

/* Warm start of one LP (extension): the point a scenario re-solve starts from,
 * usually the solution of the base LP (py/refinery/reoptimization.py:31-79
 * re-solves each perturbation from the previous solution's basis). */
typedef struct mpdhg_warm_start {
    const double *x;       /* [n] primal start, or NULL: clamp(0, l, u) */
    const double *y;       /* [m] row duals in the config's dual_sign, or NULL: 0 */
    double primal_weight;  /* enhanced: omega to start from; <= 0: ||c|| / ||b|| of the scaled LP */
    double step_size;      /* enhanced: eta; <= 0: the power iteration (every LP of the batch
                              must give one for the power iteration to be skipped) */
} mpdhg_warm_start;
/* mpdhg_solve_batch with a warm start per LP (warm may be NULL: all cold).
 * Reference mode starts from x (in FP32, clamped to the bounds) and y
 * instead of (clamp(0, l, u), 0) and keeps gpu_solver.cpp:227-271 otherwise;
 * enhanced mode starts its iterates and its Halpern anchor at the scaled x
 * and y, with the given primal weight and step size. */
int mpdhg_solve_warm(mpdhg_ctx *ctx, const mpdhg_lp *lps, int count, const mpdhg_config *cfg,
                     const mpdhg_warm_start *warm, mpdhg_result *res);

/* Adjoint of the PDHG fixed point (extension; no reference counterpart).
 * For L(x*) with upstream gradient w = dL/dx* at a solution (x, y) of `lp`:
 * columns strictly inside their bounds form B, rows with nonzero duals or
 * tight at x form R, and the transposed fixed-point condition of the projected PDHG step on
 * (B, R) reduces to A_RB^T mu = w_B, solved per instance on the GPU by CGLS
 * (min-norm least squares in the scaled space).  Outputs:
 *   dl_dc = 0 (the solution is piecewise constant in c on the active set),
 *   dl_db[i] = mu_i: derivative w.r.t. a shift of row i's bounds (rl and ru
 *              move together; only the active one matters),
 *   dl_dl[j] / dl_du[j] = w_j - (A^T mu)_j for columns held at that bound.
 * Any output pointer may be NULL. */
typedef struct mpdhg_adjoint_io {
    const double *x;   /* [n] forward primal solution */
    const double *y;   /* [m] forward row duals, reference sign (c + A^T y) */
    const double *w;   /* [n] dL/dx */
    double *dl_dc;     /* [n] */
    double *dl_db;     /* [m] */
    double *dl_dl;     /* [n] */
    double *dl_du;     /* [n] */
    int32_t basic_cols, active_rows, iterations; /* out */
    double residual;   /* out: final |K^T r| / |K^T w| of CGLS */
} mpdhg_adjoint_io;

/* active_tol: a column is basic when both bound distances exceed
 * active_tol * (1 + |x_j|); a row is active when |y_i| > active_tol *
 * (1 + max |y|) or (A x)_i is within active_tol * (1 + max finite |bound|) of
 * a finite bound (a tight row with a zero dual at a degenerate vertex).
 * tol: CGLS stopping tolerance; max_iter: CGLS iterations. */
int mpdhg_adjoint_batch(mpdhg_ctx *ctx, const mpdhg_lp *lps, int count, mpdhg_adjoint_io *io,
                        double active_tol, double tol, int max_iter);

/* Basis identification of GpuSolver::solve_hybrid (gpu_solver.cpp:319-395):
 * the m most interior of the n columns and m row activities become Basic,
 * the rest sit at the nearer bound.  Ties keep index order (the reference's
 * std::sort leaves them unspecified).  col_basis[n], row_basis[m]. */
int mpdhg_identify_basis(const mpdhg_lp *lp, const double *x, int8_t *col_basis,
                         int8_t *row_basis);

/* ---- FP64 crossover: SimplexSolver::solve_from_basis ------------------ */
/* Options of the polish stage (model::Options fields it reads,
 * options.hpp:85-91). */
typedef struct mpdhg_polish_opts {
    int64_t max_iterations;  /* iteration_limit: -1 = 10 (m+n) + 10000, as the reference */
    double primal_tol;       /* primal_feasibility_tol, 1e-6 */
    double dual_tol;         /* dual_feasibility_tol, 1e-6 */
    double obj_offset;       /* added to the reported objective (Problem::obj_offset) */
    int32_t refactor_every;  /* basis updates between refactorisations, 100 */
    double time_limit_sec;   /* 0 = none */
} mpdhg_polish_opts;
void mpdhg_polish_opts_default(mpdhg_polish_opts *opts);
/* Result of the polish (the model::Solution fields it fills).  Vector
 * outputs are caller-allocated; NULL skips one. */
typedef struct mpdhg_polish_result {
    double *x;              /* [n] */
    double *y;              /* [m] row duals, simplex sign: reduced cost = c - A^T y */
    double *reduced_costs;  /* [n] */
    double *row_activity;   /* [m] A x (model::Solution::slack) */
    int8_t *col_basis_out;  /* [n] MPDHG_BASIS_* (3 = free nonbasic) */
    int8_t *row_basis_out;  /* [m] */
    int32_t status;         /* MPDHG_STATUS_OPTIMAL / INFEASIBLE / UNBOUNDED / ITERATION_LIMIT / TIME_LIMIT */
    int32_t warm_start;     /* 1: started from the given basis; 0: the slack basis */
    int64_t iterations;     /* pivots and bound flips (simplex_iterations) */
    int64_t refactorizations;
    int64_t repaired;       /* dependent basis columns replaced by slacks */
    int64_t bound_flips;
    int64_t degenerate_pivots; /* basis changes with a zero step */
    int64_t push_steps;     /* primal push steps from x_start (not in iterations) */
    double objective;       /* obj_offset + c^T x */
    double primal_infeasibility; /* sum of basic bound violations at exit */
    double time_sec;
} mpdhg_polish_result;
/* SimplexSolver::solve_from_basis (simplex_solver.cpp:266-329), the polish
 * of GpuSolver::solve_hybrid (gpu_solver.cpp:397-417): an FP64 bounded
 * primal simplex (composite phase 1, then phase 2; Devex pricing; bound
 * perturbation) from the basis given by col_basis[n] / row_basis[m]
 * (MPDHG_BASIS_*, e.g. from mpdhg_identify_basis).  A basis without exactly
 * m basic entries, or NULL arrays, starts from the slack basis
 * (primal_simplex.cpp:79-99).  x_start[n] (may be NULL; an extension) is the
 * point the basis came from: nonbasic variables it puts inside their bounds
 * start there and a primal push moves them to bounds (crossover) before the
 * simplex.  Host code; the objective is minimised (sense is not read, as in
 * the PDHG). */
int mpdhg_polish(const mpdhg_lp *lp, const int8_t *col_basis, const int8_t *row_basis,
                 const double *x_start, const mpdhg_polish_opts *opts, mpdhg_polish_result *res);
/* GpuSolver::solve_hybrid (gpu_solver.cpp:309-417): mpdhg_solve on the
 * device, mpdhg_identify_basis on its x, mpdhg_polish from that basis (and
 * that x) and,
 * if the polish does not end optimal, mpdhg_polish from the slack basis
 * (the reference's fallback full simplex solve, :413-416; `iterations`
 * then counts both).  pdhg may be NULL (its vector outputs may be NULL too);
 * out->time_sec covers the whole call, as time_wall_sec does there. */
int mpdhg_solve_hybrid(mpdhg_ctx *ctx, const mpdhg_lp *lp, const mpdhg_config *cfg,
                       const mpdhg_polish_opts *polish_opts, mpdhg_result *pdhg,
                       mpdhg_polish_result *out);

/* ---- synthetic workloads (owned by the library) ----------------------- */
typedef struct mpdhg_lp_owned mpdhg_lp_owned;
/* Random equality LP: nnz_per_col distinct rows per column, values U[-1,1],
 * 0<=x<=10, b = A x0 with x0~U[0,10], c = A^T y0 + s, y0~N(0,1),
 * s~U[0,1] zeroed on floor(0.4 n) random columns. */
int mpdhg_gen_random(int64_t m, int64_t n, int32_t nnz_per_col, uint64_t seed,
                     mpdhg_lp_owned **out);
/* Refinery staircase: T blocks of R rows x C columns (product balances,
 * blend-quality rows, unit-capacity rows; crude, sale and inventory
 * columns; inventory links block t to t+1). */
int mpdhg_gen_staircase(int32_t T, int32_t R, int32_t C, uint64_t seed, mpdhg_lp_owned **out);
/* Scenario: copy of `base` with c_j *= exp(shock_sigma N(0,1)); if
 * zero_capacity and the LP has capacity rows, one of them gets ru = 0. */
int mpdhg_gen_scenario(const mpdhg_lp_owned *base, uint64_t seed, double shock_sigma,
                       int32_t zero_capacity, mpdhg_lp_owned **out);
/* Scenario perturbations of a staircase LP, restating the perturbations of
 * py/refinery/parameterized_generator.py:28-133 on the staircase model
 * (crude, sale and inventory columns; balance, blend and capacity rows).
 * Each multiplier is 1 + clip(N(0, std), lo, hi) with the reference's clips.
 * Costs and bounds of other LPs only take the lognormal shock. */
#define MPDHG_REGIME_NORMAL 0
#define MPDHG_REGIME_HIGH_CRUDE 1     /* crude costs x 1.30 */
#define MPDHG_REGIME_DIESEL_SURGE 2   /* sale prices x 1.30; every 4th product ("diesel") x 1.25
                                         more and its demand cap x 1.60/1.25 (the reference raises
                                         every cap by 1.25 and diesel caps by 1.60; the common 1.25
                                         is left out) */
#define MPDHG_REGIME_TURNAROUND 3     /* the first capacity row of each block x 0.65 */
#define MPDHG_REGIME_SOUR_CRUDE 4     /* crude costs x 1.10, first crude x 0.85, second x 1.25 */
#define MPDHG_REGIME_SWEET_SHORTAGE 5 /* the second crude's supply x 0.35 */
typedef struct mpdhg_scenario_opts {
    double shock_sigma;    /* every cost x exp(shock_sigma N(0,1)) (the sweep's price shock) */
    double price_std;      /* crude costs clip [-0.30, 0.40], sale prices clip [-0.25, 0.35] */
    double supply_std;     /* crude column upper bounds, clip [-0.40, 0.50] */
    double capacity_std;   /* capacity rows' upper bounds, clip [-0.25, 0.30] */
    double demand_std;     /* sale column upper bounds (demand caps), clip [-0.20, 0.25] */
    int32_t regime;        /* MPDHG_REGIME_* */
    int32_t zero_capacity; /* 1: one random capacity row gets ru = 0 (a unit outage) */
} mpdhg_scenario_opts;
void mpdhg_scenario_opts_default(mpdhg_scenario_opts *opts); /* all zero: normal regime, no change */
int mpdhg_gen_scenario_ex(const mpdhg_lp_owned *base, uint64_t seed, const mpdhg_scenario_opts *opts,
                          mpdhg_lp_owned **out);
/* Borrowed view of an owned LP (valid until mpdhg_lp_free). */
const mpdhg_lp *mpdhg_lp_view(const mpdhg_lp_owned *lp);
/* Row kinds of a staircase LP (0 balance, 1 blend, 2 capacity), else NULL. */
const int8_t *mpdhg_lp_row_kind(const mpdhg_lp_owned *lp);
void mpdhg_lp_free(mpdhg_lp_owned *lp);

/* ---- MPS / QPS files: io::read_mps / io::write_mps (io/mps_io.hpp:10-13) */
#define MPDHG_VAR_CONTINUOUS 0 /* model::VariableType (problem.hpp:19-24) */
#define MPDHG_VAR_BINARY 1
#define MPDHG_VAR_INTEGER 2
/* The model::Problem fields beside the LP data: name, sense, objective
 * constant, names, variable types and the quadratic objective 1/2 x^T Q x.
 * The PDHG path reads none of them (integrality and Q are ignored, as in
 * gpu_solver.cpp:60-307). */
typedef struct mpdhg_lp_meta {
    const char *name;
    int32_t sense;                 /* 1 minimize, -1 maximize (ObjectiveSense) */
    double obj_offset;             /* objective constant (minus the RHS of the N row) */
    const char *const *row_names;  /* [m] or NULL (the writer then names rows R0, R1, ...) */
    const char *const *col_names;  /* [n] or NULL (columns C0, C1, ...) */
    const int8_t *var_type;        /* [n] MPDHG_VAR_*, or NULL: all continuous */
    int32_t has_quadratic;         /* a QUADOBJ/QMATRIX section was read */
    int64_t q_nnz;
    const int64_t *q_col_ptr;      /* [n+1], Q in CSC (both triangles) */
    const int64_t *q_row_ind;      /* [q_nnz] */
    const double *q_vals;          /* [q_nnz] */
} mpdhg_lp_meta;
/* io::read_mps (mps_io.cpp:62-409): free- or fixed-column MPS/QPS by
 * whitespace tokens; sections NAME, OBJSENSE, ROWS, COLUMNS (with integer
 * MARKERs), RHS, RANGES, BOUNDS (UP LO FX FR MI PL BV UI LI), QUADOBJ,
 * QMATRIX, ENDATA; duplicates summed and |a_ij| <= 1e-15 dropped
 * (SparseMatrix::from_triplets, sparse_matrix.cpp:13-88).  Errors
 * (MPDHG_E_ARG, message as the reference's exception text): file not
 * readable, a malformed or out-of-range number, contradictory bounds. */
int mpdhg_read_mps(const char *path, mpdhg_lp_owned **out);
/* Borrowed view of an owned LP's meta fields (valid until mpdhg_lp_free). */
const mpdhg_lp_meta *mpdhg_lp_meta_view(const mpdhg_lp_owned *lp);
/* io::write_mps (mps_io.cpp:411-578): the same text the reference writes,
 * byte for byte (12 significant digits).  meta may be NULL (name "problem",
 * minimize, no names, continuous, no Q).  free_format is accepted and, as in
 * the reference, does not change the output. */
int mpdhg_write_mps(const char *path, const mpdhg_lp *lp, const mpdhg_lp_meta *meta,
                    int32_t free_format);

/* ---- solution checker: checker::check_solution (checker/checker.cpp:9-268) */
/* The candidate a checker call verifies (the model::Solution fields it reads).
 * Duals in the checker's sign: reduced cost = c - A^T y, the negative of the
 * PDHG sign (solve with dual_sign = 1, or negate y).  A NULL vector stands
 * for an empty one. */
typedef struct mpdhg_check_input {
    int32_t status;              /* MPDHG_STATUS_* */
    const double *x;             /* [n] */
    const double *row_duals;     /* [m] */
    const double *reduced_costs; /* [n] */
    double primal_objective;
    const double *ray;           /* Farkas ray [m] (Infeasible) or ray [n] (Unbounded) */
    int64_t ray_len;
} mpdhg_check_input;
/* checker::CheckResult (checker.hpp:9-34); summary holds the reference's text. */
typedef struct mpdhg_check_result {
    int32_t is_primal_feasible;
    double max_row_violation, max_col_bound_violation, max_primal_residual;
    int32_t is_dual_feasible;
    double max_dual_residual;
    int32_t is_complementary;
    double max_complementarity_slack;
    int32_t is_integrality_satisfied;
    double max_integrality_violation;
    double evaluated_objective, objective_discrepancy;
    int32_t is_certificate_valid;
    double certificate_violation;
    int32_t all_checks_passed;
    char summary[256];
} mpdhg_check_result;
/* Independent FP64 verification of a candidate solution of an LP, with the
 * reference's semantics and arithmetic order (A x by rows, A^T y by columns,
 * in storage order): primal row and bound violations, integrality, the
 * objective (with meta's offset and quadratic term), and for continuous
 * problems with duals the stationarity c + Qx - A^T y - s and the
 * complementarity |y_i| min(|Ax_i - rl_i|, |ru_i - Ax_i|); Farkas and
 * unbounded-ray certificates for those statuses.  meta may be NULL (no
 * offset, continuous, no Q).  Host code. */
int mpdhg_check_solution(const mpdhg_lp *lp, const mpdhg_lp_meta *meta, const mpdhg_check_input *sol,
                         double tol_primal, double tol_dual, double tol_int, mpdhg_check_result *out);

#ifdef __cplusplus
}
#endif
#endif /* MPDHG_H */
