// Types shared by the host engine and the gfx950 kernels.
#pragma once

#include <cstdint>

namespace mpdhg {

// Workgroup shape of every tile kernel: 4 wave64s, one row per lane in the
// epilogue, kTileNnz nonzeros staged through LDS per chunk.
constexpr int kBlock = 256;
constexpr int kTileNnz = 4096;
constexpr int kPerThread = kTileNnz / kBlock;

// A tile: rows [begin, end) of one instance's compressed structure
// (CSR rows for A x, CSC columns for A^T y).  end - begin <= kBlock.
struct Tile {
    int32_t begin;
    int32_t end;
    int32_t inst;
    int32_t pad;
};

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

// Enhanced-mode per-instance state.
struct EnhInst {
    int32_t status;
    int32_t iterations;       // iterations done (total, this solve)
    int32_t restarts;
    int32_t checks;
    int32_t last_restart;     // iteration index of the last restart
    int32_t restart_at;       // iteration at which K1/K2 take z+ as iterate and anchor
    int32_t row_tile_begin, row_tile_end;
    int32_t col_tile_begin, col_tile_end;
    float tau, sigma;         // eta / omega, eta * omega
    double eta, omega;
    double r_anchor, r_prev;  // fixed-point residual at last restart / previous check
    double bnorm_inf, cnorm_inf;
    double eps_p, eps_d, eps_gap, kkt, pobj, dobj;
    double tol;
    double best_kkt;
};

}  // namespace mpdhg
