"""mpdhg_solve_noalloc against mpdhg_solve, through the C ABI.

The no-allocation solve carves every device buffer from the library's
per-GPU scratch array and launches kernel by kernel with the rule's tile
schedules; neither changes a result, so its output must equal mpdhg_solve's
bit for bit (x, y, reduced costs, row activities, status, iterations, KKT) in
both modes.  Two calls in a row share the scratch, the second on a smaller
LP, so stale contents of the first call must not leak into the second.  An
LP too large for the scratch is rejected before any device work (that part
needs no GPU).
"""
import ctypes as C

import numpy as np
import pytest

import sih_solver as S
from sih_solver import _abi

SCRATCH = 64 << 20


def _call(fn, ctx, p, cfg):
    lp, keep = p._c_view()
    m, n = p.num_rows(), p.num_cols()
    out = (np.full(n, np.nan), np.full(m, np.nan), np.full(n, np.nan), np.full(m, np.nan))
    res = _abi.MpdhgResult()
    dp = C.POINTER(C.c_double)
    res.x, res.y, res.reduced_costs, res.row_activity = (a.ctypes.data_as(dp) for a in out)
    c = cfg._c()
    rc = fn(ctx, C.byref(lp), C.byref(c), C.byref(res))
    if rc != 0:
        raise _abi.MpdhgError(rc, fn.__name__)
    return res, out


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["reference", "enhanced"])
def test_noalloc_equals_solve_on_two_calls(need_gpu, mode):
    lib = _abi.lib()
    ctx = C.c_void_p()
    _abi.check(lib.mpdhg_create(0, C.byref(ctx)), "mpdhg_create")
    try:
        cfg = S.GpuPdhgConfig() if mode == "reference" else S.GpuPdhgConfig.enhanced(tol=1e-6)
        # second call smaller than the first
        for p in (S.gen_random(3000, 7200, 8, 7), S.gen_staircase(2, 100, 240, 8)):
            a, ao = _call(lib.mpdhg_solve, ctx, p, cfg)
            b, bo = _call(lib.mpdhg_solve_noalloc, ctx, p, cfg)
            assert a.status == b.status and a.iterations == b.iterations, p.name
            assert a.restarts == b.restarts and a.checks == b.checks, p.name
            for u, v in zip(ao, bo):
                assert np.array_equal(u, v), p.name
            assert a.rel_kkt == b.rel_kkt and a.primal_objective == b.primal_objective, p.name
            if mode == "reference":
                assert a.ref_primal == b.ref_primal and a.ref_dual == b.ref_dual, p.name
    finally:
        lib.mpdhg_destroy(ctx)


def test_noalloc_rejects_an_lp_beyond_the_scratch():
    # m = 200,000, n = 480,000, 8 nonzeros per column: 3.84M nonzeros need
    # well over the 64 MiB scratch; the size check comes before the context
    # and before any device work, so it holds with or without a GPU.
    m, n, k = 200_000, 480_000, 8
    lp = _abi.MpdhgLp()
    lp.m, lp.n, lp.nnz = m, n, n * k
    res = _abi.MpdhgResult()
    for mode in (_abi.MPDHG_MODE_REFERENCE, _abi.MPDHG_MODE_ENHANCED):
        c = S.GpuPdhgConfig(mode=mode)._c()
        rc = _abi.lib().mpdhg_solve_noalloc(None, C.byref(lp), C.byref(c), C.byref(res))
        assert rc == -1  # MPDHG_E_ARG
        msg = _abi.lib().mpdhg_last_error().decode()
        assert "too large for the no-allocation solve" in msg and str(SCRATCH) in msg
