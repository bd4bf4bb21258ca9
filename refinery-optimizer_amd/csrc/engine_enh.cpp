// Enhanced-mode driver (restarted reflected-Halpern PDHG with FP64 checks).
#include "engine.hpp"

namespace mpdhg {

void solve_enhanced(Context &, const mpdhg_lp *, int, const mpdhg_config &, mpdhg_result *) {
    throw Error(MPDHG_E_STATE, "enhanced mode is not available in this build");
}

}  // namespace mpdhg
