// Internal declarations shared by the host translation units of libmpdhg.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mpdhg.h"

namespace mpdhg {

// A library-owned LP (synthetic workloads).
struct OwnedLp {
    int64_t m = 0, n = 0;
    std::vector<int64_t> col_ptr, row_ind;
    std::vector<double> vals, c, l, u, rl, ru;
    std::vector<int8_t> row_kind;  // staircase only
    mpdhg_lp view{};
};

OwnedLp *gen_random(int64_t m, int64_t n, int32_t k, uint64_t seed);
OwnedLp *gen_staircase(int32_t T, int32_t R, int32_t C, uint64_t seed);
OwnedLp *gen_scenario(const OwnedLp *base, uint64_t seed, double shock, int32_t zero_cap);

// Records a thread-local message and returns `code`.
int fail(int code, const std::string &msg);

}  // namespace mpdhg
