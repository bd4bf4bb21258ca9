"""The bench-only rocSPARSE baseline (include/mpdhg_vendor.h) on the GPU.

It runs the reference iteration with rocSPARSE's summation order, so it is
compared with the oracle to FP32-rounding tolerance, not bit for bit: after
K iterations of the (nonexpansive) PDHG map, max |x_vendor - x_oracle| <=
1e-3 max(1, |x|) and the objectives agree to 1e-4 relative (x_last: the
oracle's final primal iterate).  The baseline runs in a child process
because other tests of the session import torch, whose bundled rocSPARSE
would otherwise be bound in place of the system one.
"""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import sih_solver as S
from sih_solver import vendor

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent

CHILD = """
import json, sys
import numpy as np
sys.path.insert(0, {pkg!r})
import sih_solver as S
from sih_solver import vendor
p = S.gen_random(1000, 2500, 8, 0)
v = vendor.run(p, {K}, {mode})
np.save({out!r}, v.x)
print(json.dumps(dict(us_per_iter=v.us_per_iter, us_spmv=v.us_spmv, us_spmv_t=v.us_spmv_t,
                      objective=v.objective)))
"""


@pytest.mark.parametrize("mode", [vendor.TRANSPOSE_OP, vendor.EXPLICIT_CSC])
def test_vendor_baseline_tracks_the_oracle(need_gpu, mode, tmp_path):
    p = S.gen_random(1000, 2500, 8, 0)
    K = 300
    out = tmp_path / "x.npy"
    code = CHILD.format(pkg=str(ROOT / "refinery-optimizer_amd"), K=K, mode=mode, out=str(out))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    v = json.loads(r.stdout.strip().splitlines()[-1])
    x = np.load(out)
    ref = O.pdhg_ref(O.LpArrays.of(p), max_iterations=K, tol_primal=np.float32(0.0),
                     tol_dual=np.float32(0.0))
    xl = ref["x_last"].astype(np.float64)
    scale = max(1.0, np.abs(xl).max())
    assert np.abs(x - xl).max() <= 1e-3 * scale
    obj = float(p.c @ xl)
    assert abs(v["objective"] - obj) <= 1e-4 * max(1.0, abs(obj))
    assert v["us_per_iter"] > 0 and v["us_spmv"] > 0 and v["us_spmv_t"] > 0


MEASURE = """
import json, sys
sys.path.insert(0, {pkg!r})
from sih_solver import vendor
g = vendor.gather(2.0, per_xcd=16, reps=3)
s = vendor.stream(1 << 24, 2)
print(json.dumps(dict(gather=g, stream=s)))
"""


def test_measurement_kernels_run(need_gpu):
    # the streaming and gather ceilings the roofline fractions are read against
    code = MEASURE.format(pkg=str(ROOT / "refinery-optimizer_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    v = json.loads(r.stdout.strip().splitlines()[-1])
    g = v["gather"]
    assert g["entries"] == 8 * 16 * 8192
    assert g["stream_us"] > 0 and g["gather_us"] > 0 and g["staged_us"] > 0
    assert abs(g["stream_gbs"] - 8 * g["entries"] / (1e-6 * g["stream_us"]) / 1e9) < 1e-6 * g["stream_gbs"]
    assert min(v["stream"].values()) > 0
