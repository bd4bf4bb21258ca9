"""The bench-only rocSPARSE baseline (include/mpdhg_vendor.h) on the GPU.

It runs the reference iteration with rocSPARSE's summation order, so it is
compared with the oracle to FP32-rounding tolerance, not bit for bit: after
K iterations of the (nonexpansive) PDHG map, max |x_vendor - x_oracle| <=
1e-3 max(1, |x|) and the objectives agree to 1e-4 relative (x_last: the
oracle's final primal iterate).
"""
import numpy as np
import pytest

import oracle as O
import sih_solver as S
from sih_solver import vendor

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [vendor.TRANSPOSE_OP, vendor.EXPLICIT_CSC])
def test_vendor_baseline_tracks_the_oracle(need_gpu, mode):
    p = S.gen_random(1000, 2500, 8, 0)
    K = 300
    ref = O.pdhg_ref(O.LpArrays.of(p), max_iterations=K, tol_primal=np.float32(0.0),
                     tol_dual=np.float32(0.0))
    xl = ref["x_last"].astype(np.float64)
    v = vendor.run(p, K, mode)
    scale = max(1.0, np.abs(xl).max())
    assert np.abs(v.x - xl).max() <= 1e-3 * scale
    obj = float(p.c @ xl)
    assert abs(v.objective - obj) <= 1e-4 * max(1.0, abs(obj))
    assert v.us_per_iter > 0 and v.us_spmv > 0 and v.us_spmv_t > 0
