"""Reference mode on the GPU against the oracle: bit-identical iterates.

The engine's reference mode restates gpu_solver.cpp:60-307 / cuda_pdhg.cu:14-151
with the same per-row FP32 operation order, so x_best, the final y, the
iteration counter, the status and the reference residuals P and D must equal
the oracle's bit for bit (the tolerance for this floating-point path is zero).
"""
import math

import numpy as np
import pytest

import oracle as O
import sih_solver as S
from lpfix import random_lp, toy01

pytestmark = pytest.mark.gpu


def cfg(**kw):
    c = S.GpuPdhgConfig()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def oracle_of(p, c):
    return O.pdhg_ref(O.LpArrays.of(p), max_iterations=c.max_iterations,
                      tol_primal=np.float32(c.tol_primal), tol_dual=np.float32(c.tol_dual),
                      check_frequency=c.check_frequency, step_tau=np.float32(c.step_tau),
                      step_sigma=np.float32(c.step_sigma))


def assert_bitexact(sol, r, label=""):
    assert int(sol.status) == r["status"], label
    assert sol.simplex_iterations == r["iterations"], label
    assert np.array_equal(sol.x, r["x"]), label
    assert np.array_equal(sol.row_duals, r["y"]), label
    assert sol.primal_objective == r["objective"], label
    if r["checks"]:
        assert sol.ref_primal == r["last_primal"] and sol.ref_dual == r["last_dual"], label


def test_device_and_vram(need_gpu):
    # tests/test_gpu_basic.py:21-30
    assert S.GpuSolver.is_cuda_available()
    v = S.GpuSolver.get_vram_info()
    assert v.total_bytes / 2**20 > 1000.0 and v.used_bytes == v.total_bytes - v.free_bytes


def test_toy01_bitexact_and_hybrid_handoff(need_gpu):
    # tests/test_gpu_basic.py:44-53 configuration
    p = toy01()
    c = cfg(max_iterations=2000, tol_primal=1e-4)
    sol = S.solve_gpu_pdhg(p, c)
    assert_bitexact(sol, oracle_of(p, c), "toy01")
    hyb = S.solve_gpu_hybrid(p, c)
    cb, rb = O.identify_basis(O.LpArrays.of(p), hyb.x)
    assert [int(b) for b in hyb.col_basis] == cb.tolist()
    assert [int(b) for b in hyb.row_basis] == rb.tolist()


def test_random_equality_lp_m1000_bitexact(need_gpu):
    p = S.gen_random(1000, 2500, 8, 0)
    c = cfg()  # reference defaults: 4000 iterations, check every 25
    sol = S.solve_gpu_pdhg(p, c)
    assert sol.csr_bands == 1  # xbar of 10 kB: no column bands by rule
    r = oracle_of(p, c)
    assert_bitexact(sol, r, "random m=1000")
    # The A xbar product split into column bands (each row's sum carried from
    # band to band in storage order): the same bits, including bands that
    # hold no nonzero of some rows and 16 bands of ~156 columns.
    for bands in (2, 5, 16):
        sb = S.solve_gpu_pdhg(p, cfg(spmv_bands=bands))
        assert sb.csr_bands == bands
        assert_bitexact(sb, r, f"random m=1000, {bands} bands")


def _mixed_batch():
    probs = [toy01()]
    for s, (m, n, d) in enumerate([(30, 45, 0.2), (64, 200, 0.05), (1, 5, 1.0), (257, 300, 0.02)]):
        probs.append(random_lp(m, n, d, 100 + s)[0])
    probs.append(S.gen_staircase(2, 40, 90, 5))
    # a row longer than one LDS chunk (4096 nonzeros) and columns of one entry
    n = 9000
    p = S.Problem("long_row")
    p.resize(3, n)
    rows = np.concatenate([[0, 1 + (j % 2)] for j in range(n)]).astype(np.int64)
    vals = np.concatenate([[1.0 + (j % 7) * 0.1, -0.5] for j in range(n)])
    p.A = S.SparseMatrix.from_csc(3, n, np.arange(0, 2 * n + 1, 2), rows, vals)
    p.c = np.linspace(-1, 1, n)
    p.row_lower = np.array([10.0, -np.inf, -5.0])
    p.row_upper = np.array([10.0, 3.0, np.inf])
    p.col_upper = np.full(n, 4.0)
    probs.append(p)
    return probs


def test_mixed_batch_bitexact_and_equal_to_single_solves(need_gpu):
    probs = _mixed_batch()
    c = cfg(max_iterations=600, check_frequency=25, tol_primal=1e-5, tol_dual=1e-5)
    sols = S.solve_batch(probs, c)
    for p, s in zip(probs, sols):
        assert_bitexact(s, oracle_of(p, c), p.name)
    for p, s in zip(probs[:3], sols[:3]):
        one = S.solve_gpu_pdhg(p, c)
        assert np.array_equal(one.x, s.x) and np.array_equal(one.row_duals, s.row_duals)


def test_edge_configurations(need_gpu):
    p = random_lp(50, 80, 0.1, 9)[0]
    for kw in (dict(max_iterations=0), dict(max_iterations=1), dict(check_frequency=1, max_iterations=40),
               dict(step_tau=0.01, step_sigma=0.02, max_iterations=300),
               dict(max_iterations=251, check_frequency=250), dict(max_iterations=260, check_frequency=25)):
        for graph in (1, 0):  # full check cycles replayed as a recorded graph, or launched one by one
            c = cfg(launch_graph=graph, **kw)
            assert_bitexact(S.solve_gpu_pdhg(p, c), oracle_of(p, c), f"{kw} graph={graph}")
    z = S.solve_gpu_pdhg(p, cfg(max_iterations=0))
    assert z.simplex_iterations == 1 and z.status == S.SolutionStatus.IterationLimit
    assert np.array_equal(z.x, np.clip(0.0, p.col_lower, p.col_upper))


def test_empty_problems(need_gpu):
    p = S.Problem("empty")
    p.resize(0, 3)
    sol = S.solve_gpu_pdhg(p)
    assert sol.status == S.SolutionStatus.Optimal  # gpu_solver.cpp:77-80


def test_dual_sign_option_and_kkt_fields(need_gpu):
    p = random_lp(40, 60, 0.15, 4)[0]
    a = S.solve_gpu_pdhg(p, cfg(max_iterations=500))
    b = S.solve_gpu_pdhg(p, cfg(max_iterations=500, dual_sign=1))
    assert np.array_equal(a.row_duals, -b.row_duals)
    k = O.kkt(O.LpArrays.of(p), a.x, a.row_duals)
    assert math.isclose(a.rel_kkt, k["kkt"], rel_tol=1e-9, abs_tol=1e-15)
    np.testing.assert_allclose(a.reduced_costs, p.c + p.A.mat_trans_vec(a.row_duals), rtol=1e-12,
                               atol=1e-12)
    np.testing.assert_allclose(a.slack, p.A.mat_vec(a.x), rtol=1e-12, atol=1e-12)


SCHEDULES = [(128, 1024), (128, 2048), (256, 1024), (256, 2048), (256, 4096), (512, 4096), (512, 8192)]


def test_layouts_and_schedules_are_bit_identical(need_gpu):
    # spmv_layout 0 stores each chunk sorted by gathered index, 1 keeps
    # storage order; (spmv_block, spmv_chunk) sets the rows per workgroup and
    # the nonzeros staged per pass (and so the tiling).  The per-row summation
    # order is the same in every case and the reference check reduces with
    # max, so every result equals the oracle's bit for bit.
    probs = _mixed_batch()
    c = cfg(max_iterations=300, check_frequency=25)
    refs = [oracle_of(p, c) for p in probs]
    for layout in (0, 1):
        for block, chunk in [(0, 0)] + SCHEDULES:
            c.spmv_layout, c.spmv_block, c.spmv_chunk = layout, block, chunk
            for p, s, r in zip(probs, S.solve_batch(probs, c), refs):
                assert_bitexact(s, r, f"{p.name} layout={layout} block={block} chunk={chunk}")
                if block:
                    assert (s.csr_block, s.csr_chunk, s.csc_block, s.csc_chunk) == (block, chunk, block, chunk)
    # the schedule search (every schedule timed on this batch) picks one of them
    c.spmv_layout, c.spmv_block, c.spmv_chunk, c.spmv_tune = 0, 0, 0, 2
    sols = S.solve_batch(probs, c)
    assert sols[0].tune_ms > 0.0 and (sols[0].csr_block, sols[0].csr_chunk) in SCHEDULES
    for p, s, r in zip(probs, sols, refs):
        assert_bitexact(s, r, f"{p.name} searched schedule")
