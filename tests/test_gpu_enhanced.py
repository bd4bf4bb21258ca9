"""Enhanced mode on the GPU against its CPU port and against HiGHS.

Tolerances:
  * first check cycle (64 iterations, no restart decision yet): x and y equal
    the CPU port bit for bit when both use the same step size eta;
  * whole solves (restarts, primal weights, refinement rounds, the stop):
    every output equals the CPU port bit for bit;
  * a batch equals the same LPs solved one by one, bit for bit;
  * converged solutions: the FP64 relative KKT recomputed by the oracle from
    the returned (x, y) is <= the requested tolerance, and the objective
    matches HiGHS to 10 x that tolerance (relative).
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import sih_solver as S
from lpfix import random_lp, toy01
from test_oracle import _highs

pytestmark = pytest.mark.gpu


def enh(**kw):
    return S.GpuPdhgConfig.enhanced(**kw)


def port(p, sol, **kw):
    return O.pdhg_enh(O.LpArrays.of(p), step=sol.step_size, **kw)


def test_first_cycle_bitexact_vs_cpu_port(need_gpu):
    for p in (S.gen_random(300, 720, 8, 3), S.gen_staircase(2, 40, 90, 4), random_lp(60, 90, 0.1, 2)[0]):
        sol = S.solve_gpu_pdhg(p, enh(tol=0.0, max_iterations=64, max_refinements=0))
        ref = port(p, sol, tol=0.0, max_iterations=64, check_frequency=64, max_refinements=0)
        assert sol.simplex_iterations == ref["iterations"] == 64, p.name
        assert np.array_equal(sol.x, ref["x"]), p.name
        assert np.array_equal(sol.row_duals, ref["y"]), p.name
        assert math.isclose(sol.rel_kkt, ref["kkt"], rel_tol=1e-9), p.name
        free = O.pdhg_enh(O.LpArrays.of(p), tol=0.0, max_iterations=64, max_refinements=0)
        assert math.isclose(sol.step_size, free["step"], rel_tol=1e-5), p.name  # power iteration


def test_batch_equals_single_solves(need_gpu):
    probs = [S.gen_random(200, 480, 8, 11), S.gen_staircase(1, 200, 480, 12), toy01(),
             random_lp(80, 120, 0.08, 13)[0]]
    cfg = enh(tol=1e-6)
    sols = S.solve_batch(probs, cfg)
    for p, s in zip(probs, sols):
        one = S.solve_gpu_pdhg(p, cfg)
        assert one.simplex_iterations == s.simplex_iterations, p.name
        assert np.array_equal(one.x, s.x) and np.array_equal(one.row_duals, s.row_duals), p.name


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_converges_and_matches_highs(need_gpu, tol):
    probs = [S.gen_random(300, 720, 8, 3), S.gen_random(1000, 2500, 8, 0),
             S.gen_staircase(2, 400, 900, 1), toy01(), random_lp(120, 200, 0.05, 21)[0]]
    sols = S.solve_batch(probs, enh(tol=tol, max_iterations=200000))
    for p, s in zip(probs, sols):
        lp = O.LpArrays.of(p)
        assert s.status == S.SolutionStatus.Optimal, (p.name, s.rel_kkt)
        k = O.kkt(lp, s.x, s.row_duals)
        assert k["kkt"] <= tol * 1.0001, (p.name, k)
        assert math.isclose(k["kkt"], s.rel_kkt, rel_tol=1e-6, abs_tol=1e-12), p.name
        h, _ = _highs(lp)
        assert abs(s.primal_objective - h.fun) <= 10 * tol * (1 + abs(h.fun)), (p.name, s.primal_objective, h.fun)


def test_staircase_12_blocks_to_1e8_matches_highs_fixture(need_gpu):
    # The refinery staircase (12 blocks of 400 x 900, m = 4800) to 1e-8; its
    # FP64 HiGHS objective is a committed fixture (tests/golden/make_highs_fixtures.py).
    import json
    from pathlib import Path

    fx = json.loads((Path(__file__).parent / "golden" / "highs_staircase.json").read_text())
    ref = next(c for c in fx if (c["T"], c["R"], c["C"], c["seed"]) == (12, 400, 900, 1))
    p = S.gen_staircase(12, 400, 900, 1)
    assert (p.num_rows(), p.num_cols(), p.num_nonzeros()) == (ref["m"], ref["n"], ref["nnz"])
    s = S.solve_gpu_pdhg(p, enh(tol=1e-8, max_iterations=400000))
    assert s.status == S.SolutionStatus.Optimal and s.rel_kkt <= 1e-8
    assert O.kkt(O.LpArrays.of(p), s.x, s.row_duals)["kkt"] <= 1e-8 * 1.0001
    assert abs(s.primal_objective - ref["objective"]) <= 1e-7 * abs(ref["objective"])


FULL_CASES = {
    "staircase2_1e-8": (lambda: S.gen_staircase(2, 400, 900, 1), 1e-8),
    "random300_1e-6": (lambda: S.gen_random(300, 720, 8, 3), 1e-6),
    "random_lp_1e-7": (lambda: random_lp(120, 200, 0.05, 21)[0], 1e-7),
    "toy01_1e-8": (toy01, 1e-8),
    # a positive target no solve reaches (tol <= 0 selects the reference's
    # tolerances, mpdhg_config.tol_kkt)
    "iteration_limit": (lambda: S.gen_random(1000, 2500, 8, 0), 1e-300),
}


@pytest.mark.parametrize("case", sorted(FULL_CASES))
def test_full_solve_bitexact_vs_cpu_port(need_gpu, case):
    # The whole solve -- power-iteration step, every restart and primal-weight
    # decision, every refinement round, the stop -- equals the CPU port bit
    # for bit: the port forms each FP64 decision quantity in the engine's
    # fixed order (tests/test_oracle.py pins that order).
    make, tol = FULL_CASES[case]
    p = make()
    K = 3000 if tol < 1e-100 else 400000
    sol = S.solve_gpu_pdhg(p, enh(tol=tol, max_iterations=K))
    ref = O.pdhg_enh(O.LpArrays.of(p), tol=tol, max_iterations=K, check_frequency=64, max_refinements=8,
                     refine_ratio=1e-2)
    assert sol.step_size == ref["step"], case
    got = (int(sol.status), sol.simplex_iterations, sol.restarts, sol.refinements)
    assert got == (ref["status"], ref["iterations"], ref["restarts"], ref["refinements"]), case
    assert np.array_equal(sol.x, ref["x"]) and np.array_equal(sol.row_duals, ref["y"]), case
    assert (sol.rel_kkt, sol.rel_primal, sol.rel_dual, sol.rel_gap) == (
        ref["kkt"], ref["eps_p"], ref["eps_d"], ref["eps_gap"]), case
    assert (sol.primal_objective, sol.dual_bound) == (ref["primal_obj"], ref["dual_obj"]), case
    if K > 3000:
        assert sol.status == S.SolutionStatus.Optimal and sol.rel_kkt <= tol, case
    else:
        assert sol.status == S.SolutionStatus.IterationLimit and sol.simplex_iterations == K, case


def test_nan_cost_reports_numerical_failure(need_gpu):
    p = S.gen_random(50, 120, 4, 5)
    p.c = p.c.copy()
    p.c[7] = math.nan
    sol = S.solve_gpu_pdhg(p, enh(tol=1e-6, max_iterations=1000))
    assert sol.status == S.SolutionStatus.NumericalFailure


def test_iteration_limit_and_short_runs(need_gpu):
    p = S.gen_random(100, 240, 8, 6)
    for K in (1, 10, 64, 65, 130):
        s = S.solve_gpu_pdhg(p, enh(tol=1e-12, max_iterations=K, max_refinements=0))
        assert s.status == S.SolutionStatus.IterationLimit and s.simplex_iterations == K
        k = O.kkt(O.LpArrays.of(p), s.x, s.row_duals)
        assert math.isclose(k["kkt"], s.rel_kkt, rel_tol=1e-6), K
    e = S.Problem("empty")
    e.resize(3, 0)
    assert S.solve_gpu_pdhg(e, enh()).status == S.SolutionStatus.Optimal


@pytest.mark.parametrize("layout,block,chunk", [(1, 0, 0), (0, 128, 1024), (0, 128, 2048), (0, 256, 1024),
                                                (0, 256, 2048), (1, 256, 4096), (0, 512, 4096), (0, 512, 8192),
                                                (1, 512, 8192)])
def test_spmv_schedules_are_bit_identical(need_gpu, layout, block, chunk):
    # The tile schedule (layout, workgroup size, chunk size) changes how rows
    # are grouped into workgroups, never a result: row sums keep storage
    # order, and the FP64 sums behind step size, restarts and stopping are
    # formed per 16-row group and per instance in a fixed order.  Same
    # iterates, decisions and KKT bits.
    probs = [S.gen_random(300, 720, 8, 3), S.gen_staircase(2, 40, 90, 4), random_lp(60, 90, 0.1, 2)[0],
             toy01(), S.gen_random(2000, 4800, 8, 9)]
    a = S.solve_batch(probs, enh(tol=1e-6))
    cfg = enh(tol=1e-6)
    cfg.spmv_layout, cfg.spmv_block, cfg.spmv_chunk = layout, block, chunk
    b = S.solve_batch(probs, cfg)
    if block:
        assert (b[0].csr_block, b[0].csr_chunk, b[0].csc_block, b[0].csc_chunk) == (block, chunk, block, chunk)
    for p, s, t in zip(probs, a, b):
        assert s.simplex_iterations == t.simplex_iterations and s.restarts == t.restarts, p.name
        assert np.array_equal(s.x, t.x) and np.array_equal(s.row_duals, t.row_duals), p.name
        assert s.rel_kkt == t.rel_kkt and s.step_size == t.step_size, p.name


def test_schedule_search_changes_speed_only(need_gpu):
    # spmv_tune = 2 times every compiled schedule on the batch's own matrix
    # (layouts built on the device) and keeps the fastest per direction;
    # whatever it picks, the results equal those of the rule's schedule.
    probs = [S.gen_random(300, 720, 8, 3), S.gen_staircase(2, 40, 90, 4), toy01(),
             S.gen_random(2000, 4800, 8, 9), S.gen_staircase(3, 200, 450, 5)]
    cfg = enh(tol=1e-6)
    cfg.spmv_tune = 0
    a = S.solve_batch(probs, cfg)
    cfg.spmv_tune = 2
    b = S.solve_batch(probs, cfg)
    schedules = {(128, 1024), (128, 2048), (256, 1024), (256, 2048), (256, 4096), (512, 4096), (512, 8192)}
    assert a[0].tune_ms == 0.0 and b[0].tune_ms > 0.0
    assert (b[0].csr_block, b[0].csr_chunk) in schedules and (b[0].csc_block, b[0].csc_chunk) in schedules
    for p, s, t in zip(probs, a, b):
        assert s.simplex_iterations == t.simplex_iterations and s.restarts == t.restarts, p.name
        assert np.array_equal(s.x, t.x) and np.array_equal(s.row_duals, t.row_duals), p.name
        assert s.rel_kkt == t.rel_kkt, p.name


def test_graph_replay_equals_eager_launches(need_gpu):
    # Full check cycles replay a recorded HIP graph, re-recorded after each
    # tile-list compaction; eager work (status copies, compactions) runs
    # between replays.  Every result must equal kernel-by-kernel launches.
    probs = [S.gen_random(200, 480, 8, 31), S.gen_staircase(1, 200, 480, 32), toy01(),
             random_lp(80, 120, 0.08, 33)[0], S.gen_random(600, 1440, 8, 34),
             S.gen_staircase(2, 100, 240, 35)]
    cfg = enh(tol=1e-6)
    g = S.solve_batch(probs, cfg)
    cfg.launch_graph = 0
    e = S.solve_batch(probs, cfg)
    assert len({s.simplex_iterations for s in g}) > 2  # instances stop in different cycles
    for p, s, t in zip(probs, g, e):
        assert s.simplex_iterations == t.simplex_iterations and s.restarts == t.restarts, p.name
        assert np.array_equal(s.x, t.x) and np.array_equal(s.row_duals, t.row_duals), p.name
        assert s.rel_kkt == t.rel_kkt and s.status == t.status, p.name


def _edge_lp():
    """random_lp(60, 100, 0.1, 1) widened by an empty row (bounds [-1, 1]) and
    two empty columns (cost 0.7 on [0, 5], cost -0.4 on [-2, 3]), with 8 of
    its columns made free.  HiGHS finds it bounded and feasible."""
    base, dense = random_lp(60, 100, 0.1, 1)
    m, n = 61, 102
    a = np.zeros((m, n))
    a[:60, :100] = dense
    free = np.random.default_rng(101).choice(100, 8, replace=False)
    l = np.concatenate([base.col_lower, [0.0, -2.0]])
    u = np.concatenate([base.col_upper, [5.0, 3.0]])
    l[free], u[free] = -math.inf, math.inf
    ptr, rows, vals = [0], [], []
    for j in range(n):
        nz = np.nonzero(a[:, j])[0]
        rows += nz.tolist()
        vals += a[nz, j].tolist()
        ptr.append(len(rows))
    p = S.Problem("edge_free_empty")
    p.resize(m, n)
    p.A = S.SparseMatrix.from_csc(m, n, np.array(ptr), np.array(rows, np.int64), np.array(vals))
    p.c = np.concatenate([base.c, [0.7, -0.4]])
    p.col_lower, p.col_upper = l, u
    p.row_lower = np.concatenate([base.row_lower, [-1.0]])
    p.row_upper = np.concatenate([base.row_upper, [1.0]])
    return p


def test_edge_instances_in_one_batch(need_gpu):
    # Instances with m = 0 or n = 0 are resolved on the host (Optimal, no
    # iteration, as gpu_solver.cpp:77-80 does); the others iterate on the
    # device in the same batch.  The edge LP's first check cycle equals the
    # CPU port bit for bit, it converges to HiGHS's optimum, and every
    # instance of the batch equals its single solve.
    edge = _edge_lp()
    cols_only, rows_only = S.Problem("m0"), S.Problem("n0")
    cols_only.resize(0, 4)
    rows_only.resize(3, 0)
    probs = [cols_only, edge, rows_only, S.gen_random(200, 480, 8, 41)]

    sol = S.solve_gpu_pdhg(edge, enh(tol=0.0, max_iterations=64, max_refinements=0))
    ref = port(edge, sol, tol=0.0, max_iterations=64, check_frequency=64, max_refinements=0)
    assert np.array_equal(sol.x, ref["x"]) and np.array_equal(sol.row_duals, ref["y"])

    cfg = enh(tol=1e-6, max_iterations=200000)
    sols = S.solve_batch(probs, cfg)
    assert sols[0].status == sols[2].status == S.SolutionStatus.Optimal
    assert sols[0].simplex_iterations == sols[2].simplex_iterations == 0
    lp = O.LpArrays.of(edge)
    s = sols[1]
    assert s.status == S.SolutionStatus.Optimal
    assert O.kkt(lp, s.x, s.row_duals)["kkt"] <= 1e-6 * 1.0001
    assert s.x[100] == 0.0 and s.x[101] == 3.0  # empty columns sit at their cost's bound
    h, _ = _highs(lp)
    assert abs(s.primal_objective - h.fun) <= 1e-5 * (1 + abs(h.fun))
    for p, t in zip(probs[1:], sols[1:]):
        if p.num_rows() and p.num_cols():
            one = S.solve_gpu_pdhg(p, cfg)
            assert one.simplex_iterations == t.simplex_iterations, p.name
            assert np.array_equal(one.x, t.x) and np.array_equal(one.row_duals, t.row_duals), p.name


def test_stream_cache_policy_changes_speed_only(need_gpu):
    # A batch whose matrix stream exceeds the last-level cache (256 MiB per
    # iteration, both directions) runs the fused kernels that load the stream
    # non-temporally; one LP alone runs the cached kernels.  Only the cache
    # policy differs, so every LP of the batch equals its single solve.
    big = S.gen_random(1_000_000, 2_400_000, 8, 51)  # 19.2M nonzeros: 307 MB of stream
    small = [S.gen_random(300, 720, 8, 52), S.gen_staircase(2, 40, 90, 53), toy01()]
    cfg = enh(tol=1e-6, max_iterations=192)
    sols = S.solve_batch([big] + small, cfg)
    assert 16.0 * big.num_nonzeros() > 256 * 2**20
    for p, s in zip(small, sols[1:]):
        one = S.solve_gpu_pdhg(p, cfg)
        assert one.simplex_iterations == s.simplex_iterations and one.restarts == s.restarts, p.name
        assert np.array_equal(one.x, s.x) and np.array_equal(one.row_duals, s.row_duals), p.name
        assert one.rel_kkt == s.rel_kkt, p.name


@pytest.mark.parametrize("bands", [2, 5, 16])
def test_column_bands_are_bit_identical(need_gpu, bands):
    # The fused dual kernel's A xbar product split into column bands, each row
    # continuing its FMA chain from the running sum the previous band stored:
    # every iterate, decision and KKT bit equals the unbanded solve, with the
    # check cycles replayed as graphs and launched kernel by kernel.
    for p in (S.gen_random(1000, 2500, 8, 61), S.gen_staircase(3, 200, 450, 62)):
        a = S.solve_gpu_pdhg(p, enh(tol=1e-6, spmv_bands=1))
        assert a.csr_bands == 1
        for graph in (1, 0):
            b = S.solve_gpu_pdhg(p, enh(tol=1e-6, spmv_bands=bands, launch_graph=graph))
            assert b.csr_bands == bands
            assert b.simplex_iterations == a.simplex_iterations and b.restarts == a.restarts, p.name
            assert np.array_equal(a.x, b.x) and np.array_equal(a.row_duals, b.row_duals), p.name
            assert a.rel_kkt == b.rel_kkt and a.status == b.status, p.name


def test_column_bands_by_rule_on_a_large_lp(need_gpu):
    # 2.4M columns (xbar of 9.6 MB): four bands by rule, and a matrix stream of
    # 307 MB per iteration, so the banded kernels read it non-temporally.
    big = S.gen_random(1_000_000, 2_400_000, 8, 51)
    cfg = enh(tol=1e-6, max_iterations=192)
    a = S.solve_gpu_pdhg(big, cfg)
    cfg.spmv_bands = 1
    b = S.solve_gpu_pdhg(big, cfg)
    assert a.csr_bands == 4 and b.csr_bands == 1
    assert a.simplex_iterations == b.simplex_iterations == 192
    assert np.array_equal(a.x, b.x) and np.array_equal(a.row_duals, b.row_duals)
    assert a.rel_kkt == b.rel_kkt


def _long_row_lp(seed=81):
    """An LP with split rows (kSplitRow = 8192): a random equality LP of
    m = 12,000, n = 30,000 with 3 nonzeros per column, plus rows 0, 1 and 2
    over 20,000, 9,000 and 8,193 distinct columns and column 0 over 12,000
    distinct rows (values U[-1, 1]), b = A x0 for x0 ~ U[0, 10], 0 <= x <= 10,
    c = A^T y0 + s with s ~ U[0, 1]."""
    rng = np.random.default_rng(seed)
    m, n = 12000, 30000
    rows = [rng.choice(m, 3, replace=False) for _ in range(n)]
    for r, cnt in ((0, 20000), (1, 9000), (2, 8193)):
        for j in rng.choice(n, cnt, replace=False):
            rows[j] = np.union1d(rows[j], [r])
    rows[0] = np.union1d(rows[0], rng.choice(m, 12000, replace=False))
    ptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    ri = np.concatenate(rows).astype(np.int64)
    vals = rng.uniform(-1, 1, len(ri))
    p = S.Problem("long_rows")
    p.resize(m, n)
    p.A = S.SparseMatrix.from_csc(m, n, ptr, ri, vals)
    x0 = rng.uniform(0, 10, n)
    y0 = rng.normal(size=m)
    A = sp.csc_matrix((vals, ri, ptr), shape=(m, n))
    b = A @ x0
    p.c = A.T @ y0 + rng.uniform(0, 1, n)
    p.row_lower, p.row_upper = b.copy(), b.copy()
    p.col_lower, p.col_upper = np.zeros(n), np.full(n, 10.0)
    return p


@pytest.mark.parametrize("tol,K", [(1e-4, 60000), (1e-300, 2000)])
def test_split_rows_full_solve_bitexact_vs_cpu_port(need_gpu, tol, K):
    # whole solves with split rows in the FP32 products and the FP64 checks:
    # to 1e-4 (a refinement round, then the stop) and 2,000 iterations
    # (restarts, a refinement, the iteration limit)
    p = _long_row_lp()
    sol = S.solve_gpu_pdhg(p, enh(tol=tol, max_iterations=K))
    ref = O.pdhg_enh(O.LpArrays.of(p), tol=tol, max_iterations=K)
    assert sol.step_size == ref["step"]
    got = (int(sol.status), sol.simplex_iterations, sol.restarts, sol.refinements)
    assert got == (ref["status"], ref["iterations"], ref["restarts"], ref["refinements"])
    assert ref["refinements"] >= 1
    assert np.array_equal(sol.x, ref["x"]) and np.array_equal(sol.row_duals, ref["y"])
    assert (sol.rel_kkt, sol.primal_objective, sol.dual_bound) == (ref["kkt"], ref["primal_obj"], ref["dual_obj"])


def test_split_rows_first_cycle_bitexact_vs_cpu_port(need_gpu):
    # Rows and columns of more than 8192 nonzeros are summed in 64-nonzero
    # segments within 4096-nonzero blocks, by the enhanced kernels (block sums
    # from the pre-pass over many workgroups) and by the CPU port alike.
    p = _long_row_lp()
    sol = S.solve_gpu_pdhg(p, enh(tol=0.0, max_iterations=64, max_refinements=0))
    ref = port(p, sol, tol=0.0, max_iterations=64, check_frequency=64, max_refinements=0)
    assert np.array_equal(sol.x, ref["x"]) and np.array_equal(sol.row_duals, ref["y"])
    assert sol.csr_bands == 1  # no column bands for an LP with split rows


@pytest.mark.parametrize("layout,block,chunk", [(0, 128, 1024), (0, 256, 2048), (1, 256, 4096), (0, 512, 8192)])
def test_split_rows_every_schedule_and_batch_bit_identical(need_gpu, layout, block, chunk):
    p = _long_row_lp()
    other = S.gen_random(2000, 4800, 8, 82)
    a = S.solve_batch([p, other], enh(tol=1e-6))
    cfg = enh(tol=1e-6)
    cfg.spmv_layout, cfg.spmv_block, cfg.spmv_chunk = layout, block, chunk
    b = S.solve_gpu_pdhg(p, cfg)
    assert b.status == S.SolutionStatus.Optimal
    assert b.simplex_iterations == a[0].simplex_iterations and b.restarts == a[0].restarts
    assert np.array_equal(a[0].x, b.x) and np.array_equal(a[0].row_duals, b.row_duals)
    assert a[0].rel_kkt == b.rel_kkt
    assert O.kkt(O.LpArrays.of(p), b.x, b.row_duals)["kkt"] <= 1e-6 * 1.0001


def test_batch_order_changes_speed_only(need_gpu):
    # The running LPs are laid out stratified by size (each XCD's eighth of
    # the tile list holds the same mix), whatever the caller's order: the
    # same LPs in reversed and in size-sorted order give every result bit for bit.
    probs = [S.gen_random(300 + 40 * k, 720 + 96 * k, 8, 90 + k) for k in range(6)]
    probs += [S.gen_staircase(2, 40 + 10 * k, 90 + 20 * k, 96 + k) for k in range(4)] + [toy01()]
    cfg = enh(tol=1e-6)
    a = S.solve_batch(probs, cfg)
    rev = S.solve_batch(probs[::-1], cfg)[::-1]
    order = sorted(range(len(probs)), key=lambda q: probs[q].num_nonzeros())
    srt = S.solve_batch([probs[q] for q in order], cfg)
    by_q = {q: s for q, s in zip(order, srt)}
    for q, p in enumerate(probs):
        for t in (rev[q], by_q[q]):
            assert t.simplex_iterations == a[q].simplex_iterations and t.restarts == a[q].restarts, p.name
            assert np.array_equal(t.x, a[q].x) and np.array_equal(t.row_duals, a[q].row_duals), p.name
            assert t.rel_kkt == a[q].rel_kkt, p.name
