// Types shared by the host engine and the gfx950 kernels.
#pragma once

#include <cstdint>

namespace mpdhg {

// Tile schedules: a tile set runs one workgroup of `block` threads per tile
// (one row per lane in the epilogue: a tile holds at most `block` rows) and
// stages `tn` nonzeros per LDS pass (a tile holds at most tn nonzeros unless
// it is a single denser row group, streamed in several passes
// [lo + j tn, lo + (j+1) tn)).  The LDS a workgroup reserves is sized by tn.
// Each (block, tn) pair below is compiled into every tile kernel; which one
// a solve uses is a performance choice only (results do not depend on it).
struct Schedule {
    int block;
    int tn;
};
constexpr Schedule kSchedules[] = {{128, 1024}, {128, 2048}, {256, 1024}, {256, 2048},
                                   {256, 4096}, {512, 4096}, {512, 8192}};
constexpr int kNumSchedules = static_cast<int>(sizeof(kSchedules) / sizeof(kSchedules[0]));
constexpr int kBlockMax = 512;
constexpr int kTileNnz = 8192;  // largest tn
constexpr int kChunkMin = 1024;

constexpr bool valid_schedule(int block, int tn) {
    for (int i = 0; i < kNumSchedules; ++i)
        if (kSchedules[i].block == block && kSchedules[i].tn == tn) return true;
    return false;
}
constexpr bool valid_block(int block) {
    for (int i = 0; i < kNumSchedules; ++i)
        if (kSchedules[i].block == block) return true;
    return false;
}

constexpr int ilog2(int v) { return v > 1 ? 1 + ilog2(v / 2) : 0; }
constexpr bool valid_chunk(int tn) {
    for (int i = 0; i < kNumSchedules; ++i)
        if (kSchedules[i].tn == tn) return true;
    return false;
}

// Gather order: within each chunk the nonzeros may be stored sorted by the
// index of the vector entry they gather (ties in storage order), so that
// the 64 gathers of one wave instruction fall on few cache lines.  The
// index word then packs (local vector index << log2 tn) | (storage
// position in the chunk), where local = index - Tile::base, which needs
// local < gather_local_limit(tn).  The lane that owns a row still reads that
// row's products back in storage order, so every sum is formed in the same
// order either way.
constexpr int64_t gather_local_limit(int tn) { return int64_t(1) << (32 - ilog2(tn)); }

// Row groups: each instance's rows (and columns) are cut into groups of
// kGroup = 16 consecutive rows from the instance's first row, the rows of
// 16 adjacent lanes.  Tiles start on group boundaries, so every FP64 sum
// that feeds a decision is formed per group by its lanes in a fixed order
// and then per instance over its groups in a fixed order (canon_sum): no
// result depends on the tile schedule.  Small groups keep tiles packed
// close to their nonzero budget.
constexpr int kGroup = 16;
static_assert(kGroup <= 64 && (kGroup & (kGroup - 1)) == 0, "groups within a wave");

// Enhanced mode's products (FP32 iterations, FP64 checks and adjoint) sum a
// row (or column) with more than kSplitRow nonzeros in two levels, counted
// from the row's first nonzero: segments of kSplitSeg nonzeros, each a
// sequential FMA chain from zero; blocks of kSplitBlockNnz nonzeros, each the
// in-order sum of its segments' sums; and the row sum, the in-order sum of
// its blocks' sums.  Every other row is one sequential FMA chain.  kSplitRow
// is at least every chunk size, so a split row always lies in a tile of one
// dense row group (more nonzeros than one chunk), and the definition depends
// on the row alone, not on the tiling.
constexpr int kSplitRow = 8192;
constexpr int kSplitSeg = 64;
constexpr int kSplitBlockNnz = 4096;

// One block of a split row for the pre-pass: nonzeros [c0, c1) of the
// direction's storage-order stream (c0 - the row's first nonzero is a
// multiple of kSplitBlockNnz), whose block sum goes to blk[slot].
struct SplitChunk {
    int32_t c0, c1, slot, inst;
};
constexpr int kSplitBlock = 256;  // threads per pre-pass workgroup

// A tile: rows [begin, end) of one instance's compressed structure
// (CSR rows for A x, CSC columns for A^T y).  end - begin <= the tile set's
// block, and
// begin is the first row of a group (grp: that group's index in the batch).
// base >= 0: the tile's chunks are in gather order, gathered vector entries
// at base + local; base == -1: storage order, index words are global.
// [lo, hi) is the tile's nonzero range (ptr[begin], ptr[end]), held in the
// record so that a kernel can start streaming the nonzeros as soon as the
// record arrives.  run is a hint refreshed after every status change
// (1 while the instance iterates): a tile with run == 0 belongs to a
// stopped instance and its workgroup exits at once.  The instance status
// stays authoritative: kernels store nothing for a stopped instance.
struct Tile {
    int32_t begin;
    int32_t end;
    int32_t inst;
    int32_t base;
    int32_t lo;
    int32_t hi;
    int32_t run;
    int32_t grp;
};
static_assert(sizeof(Tile) == 32, "one 32-byte scalar load per tile record");

constexpr int32_t kRunning = -1;  // instance status while iterating

// Reference-mode per-instance state (gpu_solver.cpp:219-225 locals).
struct RefInst {
    int32_t status;      // kRunning, then model::SolutionStatus
    int32_t iterations;  // loop counter at exit
    int32_t checks;
    int32_t copy_best;   // set by the check when max(P,D) improved
    float best_res;      // best_combined_res
    float last_p, last_d;
    float max_abs_c, max_abs_b;
    int32_t row_tile_begin, row_tile_end;
    int32_t col_tile_begin, col_tile_end;
    int32_t pad;
};

// Per-group partials of the enhanced mode, quantity-major: quantity q of
// group g at part[q * groups + g].
// Column (CSC) groups:
enum : int {
    kGcDx2 = 0,       // sum (x - x+)^2            (fused primal step, check iterations)
    kGcDxa2 = 1,      // sum (x+ - x0)^2           (distance to the anchor)
    kGcPobj = 2,      // sum c_j x_j               (FP64 check)
    kGcDobj = 3,      // sum l_j lambda_j^+ - u_j lambda_j^-
    kGcDualViol = 4,  // max |lambda - proj(lambda)|
    kGcColViol = 5,   // max bound violation of x
    kGcPow = 6,       // power iteration: sum z_j^2
    kGcCount = 7
};
// Row (CSR) groups:
enum : int {
    kGrDy2 = 0,       // sum (y - y+)^2
    kGrDya2 = 1,      // sum (y+ - y0)^2
    kGrDobj = 2,      // - sum (ru_i y_i^+ - rl_i y_i^-)
    kGrRowViol = 3,   // max row-activity violation
    kGrSignViol = 4,  // max row-dual sign violation
    kGrCount = 5
};

// Threads of the per-instance kernels (finalize, power-iteration and CGLS
// scalars): fixed, so that their reduction trees never change.
constexpr int kInstBlock = 256;

// Enhanced-mode per-instance state (device resident; written by the
// per-instance finalize kernel, read by every tile of the instance).
struct EnhInst {
    int32_t status;          // kRunning, then model::SolutionStatus
    int32_t iterations;      // iterations done when stopped
    int32_t restarts;
    int32_t checks;
    int32_t last_restart;    // iteration count at the last restart (Halpern k origin)
    int32_t restart_at;      // iteration index that takes z+ as iterate and anchor
    int32_t rounds;          // refinement rounds done
    int32_t copy_best;       // flag: this check's point becomes the best point
    int32_t refine;          // flag: build the residual LP now
    int32_t max_iter;
    int32_t row_tile_begin, row_tile_end;
    int32_t col_tile_begin, col_tile_end;
    int32_t rgrp_begin, rgrp_end;  // the instance's row / column groups
    int32_t cgrp_begin, cgrp_end;
    float tau, sigma, inv_sigma, pad0;
    double eta, omega;
    double r_anchor, r_prev;   // fixed-point residuals (restart criteria)
    double bnorm, cnorm;       // |b|_inf, |c|_inf of the original LP
    double tol;
    double alpha_p, alpha_d;   // refinement scales of the current round
    double inv_ap, inv_ad;
    double round_start_kkt;    // KKT at the last re-centring (refinement trigger)
    double best_kkt, best_eps_p, best_eps_d, best_eps_gap, best_pobj, best_dobj;
    double eps_p, eps_d, eps_gap, kkt, pobj, dobj;  // last check
    double pow_norm2;          // power iteration: |A^T A v|^2
    double pw_int;             // primal-weight controller: sum of the log-balance errors
};

// Adjoint (CGLS on the active-set system) per-instance state.
struct AdjInst {
    int32_t done;
    int32_t iters;
    int32_t rgrp_begin, rgrp_end;  // the instance's row / column groups
    int32_t cgrp_begin, cgrp_end;
    double gamma, gamma0, alpha, beta;
};

}  // namespace mpdhg
