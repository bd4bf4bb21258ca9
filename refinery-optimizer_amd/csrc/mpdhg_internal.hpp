// Internal declarations shared by the host translation units of libmpdhg.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mpdhg.h"

namespace mpdhg {

// A library-owned LP (synthetic workloads, MPS files).
struct OwnedLp {
    int64_t m = 0, n = 0;
    std::vector<int64_t> col_ptr, row_ind;
    std::vector<double> vals, c, l, u, rl, ru;
    std::vector<int8_t> row_kind;  // staircase only: 0 balance, 1 blend, 2 capacity
    std::vector<int8_t> col_kind;  // staircase only: 0 crude, 1 sale, 2 inventory
    std::vector<int32_t> row_block, col_block;  // staircase only: time block of each row / column
    mpdhg_lp view{};
    // model::Problem fields beside the LP (MPS files; defaults otherwise)
    std::string name = "problem";
    int32_t sense = 1;
    double obj_offset = 0.0;
    std::vector<std::string> row_names, col_names;
    std::vector<int8_t> var_type;
    bool has_q = false;
    std::vector<int64_t> q_col_ptr, q_row_ind;
    std::vector<double> q_vals;
    std::vector<const char *> row_name_ptr, col_name_ptr;
    mpdhg_lp_meta meta{};
};

// Points view and meta at the owned arrays (after they stop changing).
void finish_views(OwnedLp *o);
OwnedLp *read_mps(const std::string &path);  // throws std::runtime_error
void write_mps(const std::string &path, const mpdhg_lp &lp, const mpdhg_lp_meta *meta);

OwnedLp *gen_random(int64_t m, int64_t n, int32_t k, uint64_t seed);
OwnedLp *gen_staircase(int32_t T, int32_t R, int32_t C, uint64_t seed);
OwnedLp *gen_scenario(const OwnedLp *base, uint64_t seed, double shock, int32_t zero_cap);
OwnedLp *gen_scenario_ex(const OwnedLp *base, uint64_t seed, const mpdhg_scenario_opts &o);

// Records a thread-local message and returns `code`.
int fail(int code, const std::string &msg);

}  // namespace mpdhg
