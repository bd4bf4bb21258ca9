// Enhanced-mode kernels.
#include <hip/hip_runtime.h>
#include "kernels_common.hpp"
