"""Warm starts, scenario re-optimization and the per-result FP64 vertex check.

* A warm start at the cold start point (x = clamp(0, l, u), y = 0, no primal
  weight or step) reproduces the cold solve bit for bit in both modes.
* Enhanced mode warm-started from its own 1e-6 solution (x, y, primal weight,
  step size) stops at the first check with KKT <= 1e-5, with the duals given
  in either dual sign; started from negated duals it does not.
* Enhanced mode from its 1e-6 solution re-solved to 1e-7: the first restart
  refines (the start brings its weight), the solve converges, and it equals
  the CPU port bit for bit, the reported effective weight included.
* Reference mode from a given start point equals the oracle's restatement
  started there, bit for bit, in both dual signs; enhanced mode from a given
  start (a random point, or a solution with its primal weight and step)
  equals the CPU port started there, bit for bit.
* PdhgReoptimizer (reoptimization.py:80-158 with the GPU PDHG): every
  perturbation of a 4-block staircase LP solves cold and warm to the same
  objective (10 x tol relative), and the warm solves need fewer iterations
  in total.
* vertex_check: each result's objective error against the FP64 vertex the
  crossover reaches from its point, equal to an independent crossover run.
"""
import math

import numpy as np
import pytest

import oracle as O
import sih_solver as S
from lpfix import random_lp
from sih_solver.scenarios import PdhgReoptimizer

pytestmark = pytest.mark.gpu


def _cold_point(p):
    return S.WarmStart(np.clip(np.zeros(p.num_cols()), p.col_lower, p.col_upper), np.zeros(p.num_rows()))


@pytest.mark.parametrize("mode", ["reference", "enhanced"])
def test_warm_start_at_the_cold_point_is_the_cold_solve(need_gpu, mode):
    probs = [S.gen_random(300, 720, 8, 3), S.gen_staircase(2, 40, 90, 4), random_lp(60, 90, 0.1, 2)[0]]
    cfg = S.GpuPdhgConfig() if mode == "reference" else S.GpuPdhgConfig.enhanced(tol=1e-6)
    a = S.solve_batch(probs, cfg)
    b = S.solve_batch(probs, cfg, [_cold_point(p) for p in probs])
    for p, s, t in zip(probs, a, b):
        assert s.simplex_iterations == t.simplex_iterations, p.name
        assert np.array_equal(s.x, t.x) and np.array_equal(s.row_duals, t.row_duals), p.name


@pytest.mark.parametrize("dual_sign", [0, 1])
def test_enhanced_warm_start_from_its_solution(need_gpu, dual_sign):
    p = S.gen_staircase(3, 200, 450, 6)
    cfg = S.GpuPdhgConfig.enhanced(tol=1e-6, dual_sign=dual_sign)
    cold = S.solve_gpu_pdhg(p, cfg)
    assert cold.status == S.SolutionStatus.Optimal and cold.simplex_iterations > 640
    loose = S.GpuPdhgConfig.enhanced(tol=1e-5, dual_sign=dual_sign)
    w = S.solve_batch([p], loose, [cold])[0]
    assert w.status == S.SolutionStatus.Optimal and w.simplex_iterations == 64, (w.simplex_iterations, w.rel_kkt)
    y = w.row_duals if dual_sign == 0 else -w.row_duals
    assert O.kkt(O.LpArrays.of(p), w.x, y)["kkt"] <= 1e-5 * 1.0001
    flipped = S.WarmStart(cold.x, -cold.row_duals, cold.primal_weight, cold.step_size)
    f = S.solve_batch([p], loose, [flipped])[0]
    assert f.simplex_iterations > 64


@pytest.mark.parametrize("dual_sign", [0, 1])
def test_reference_warm_start_bitexact_with_the_oracle(need_gpu, dual_sign):
    # Reference mode from a given start: x clamped to the bounds in FP32 and y,
    # then gpu_solver.cpp:227-271 unchanged; the oracle restates the same start.
    p = S.gen_random(200, 500, 8, 1)
    lp = O.LpArrays.of(p)
    rng = np.random.default_rng(11)
    xs, ys = rng.uniform(-1, 11, 500), rng.normal(size=200)  # outside the bounds too
    cfg = S.GpuPdhgConfig(max_iterations=600, dual_sign=dual_sign)
    sol = S.solve_batch([p], cfg, [S.WarmStart(xs, ys if dual_sign == 0 else -ys)])[0]
    r = O.pdhg_ref(lp, max_iterations=600, tol_primal=np.float32(1e-4), tol_dual=np.float32(1e-4),
                   x0=xs, y0=ys)
    assert sol.simplex_iterations == r["iterations"] and int(sol.status) == r["status"]
    assert np.array_equal(sol.x, r["x"])
    y = sol.row_duals if dual_sign == 0 else -sol.row_duals
    assert np.array_equal(y, r["y"])
    assert sol.ref_primal == r["last_primal"] and sol.ref_dual == r["last_dual"]


@pytest.mark.parametrize("dual_sign", [0, 1])
def test_enhanced_warm_start_bitexact_with_the_cpu_port(need_gpu, dual_sign):
    # Enhanced mode from a given start -- a random point (outside the bounds
    # too; the power iteration and the |c| / |b| weight as cold) and a 1e-4
    # solution with its primal weight and step -- equals the CPU port started
    # at the same point, bit for bit, to 1e-7.
    p = S.gen_staircase(3, 200, 450, 6)
    lp = O.LpArrays.of(p)
    rng = np.random.default_rng(17)
    sg = -1.0 if dual_sign else 1.0
    cfg = S.GpuPdhgConfig.enhanced(tol=1e-7, dual_sign=dual_sign)
    first = S.solve_gpu_pdhg(p, S.GpuPdhgConfig.enhanced(tol=1e-4, dual_sign=dual_sign))
    assert first.status == S.SolutionStatus.Optimal
    starts = [(rng.uniform(-100, 1100, p.num_cols()), rng.normal(size=p.num_rows()), 0.0, 0.0),
              (first.x, sg * first.row_duals, first.primal_weight, first.step_size)]
    for xs, ys, om, eta in starts:
        sol = S.solve_batch([p], cfg, [S.WarmStart(xs, sg * ys, om, eta)])[0]
        ref = O.pdhg_enh(lp, tol=1e-7, max_iterations=100000, x0=xs, y0=ys, omega0=om, step=eta)
        got = (int(sol.status), sol.simplex_iterations, sol.restarts, sol.refinements)
        assert got == (ref["status"], ref["iterations"], ref["restarts"], ref["refinements"]), (om, got)
        assert sol.status == S.SolutionStatus.Optimal and sol.step_size == ref["step"]
        assert np.array_equal(sol.x, ref["x"]) and np.array_equal(sg * sol.row_duals, ref["y"])
        assert sol.rel_kkt == ref["kkt"]


def test_scenario_reoptimization_warm_beats_cold(need_gpu):
    tol = 1e-6
    ro = PdhgReoptimizer(T=4, config=S.GpuPdhgConfig.enhanced(tol=tol))
    res = ro.run_benchmark()
    assert len(res) == 5
    for r in res:
        assert r["cold_status"] == r["warm_status"] == "Optimal", r
        assert abs(r["objective"] - r["cold_objective"]) <= 10 * tol * (1 + abs(r["cold_objective"])), r
    assert sum(r["warm_iters"] for r in res) < sum(r["cold_iters"] for r in res)
    assert "| **Crude Price Shock**" in ro.format_results_table(res)


def test_vertex_check_states_each_results_error(need_gpu):
    tol = 1e-6
    probs = [S.gen_random(400, 960, 8, 21), S.gen_staircase(2, 200, 450, 22), random_lp(80, 120, 0.08, 23)[0],
             S.gen_random(3000, 7200, 8, 24)]
    cfg = S.GpuPdhgConfig.enhanced(tol=tol, vertex_check=1000)
    sols = S.solve_batch(probs, cfg)
    for p, s in zip(probs[:3], sols[:3]):
        assert s.vertex_status == S.SolutionStatus.Optimal, p.name
        cb, rb = S.identify_basis(p, s.x)
        v = S.polish(p, cb, rb, s.x)
        assert v.status == S.SolutionStatus.Optimal and s.vertex_objective == v.objective, p.name
        assert s.vertex_rel_error == abs(s.primal_objective - v.objective) / max(1.0, abs(v.objective))
        assert s.vertex_rel_error <= 1e-4, (p.name, s.vertex_rel_error)
    big = sols[3]  # more rows than vertex_check: not run
    assert big.vertex_status == S.SolutionStatus.Unknown and math.isnan(big.vertex_rel_error)
    plain = S.solve_batch(probs[:1], S.GpuPdhgConfig.enhanced(tol=tol))[0]
    assert np.array_equal(plain.x, sols[0].x) and math.isnan(plain.vertex_objective)


@pytest.mark.parametrize("dual_sign", [0, 1])
def test_enhanced_warm_start_to_a_tighter_tolerance_bitexact(need_gpu, dual_sign):
    # From its own 1e-6 solution (x, y, effective primal weight, step) to
    # 1e-7: the first restart re-centres the start in FP64, so the solve
    # refines and converges; the engine equals the CPU port started at the
    # same point bit for bit, its reported primal weight included.
    p = S.gen_staircase(4, seed=3)
    lp = O.LpArrays.of(p)
    sg = -1.0 if dual_sign else 1.0
    first = S.solve_gpu_pdhg(p, S.GpuPdhgConfig.enhanced(tol=1e-6, dual_sign=dual_sign))
    ref1 = O.pdhg_enh(lp, tol=1e-6)
    assert first.status == S.SolutionStatus.Optimal and first.primal_weight == ref1["omega"]
    ws = S.WarmStart.of(first)
    sol = S.solve_batch([p], S.GpuPdhgConfig.enhanced(tol=1e-7, dual_sign=dual_sign), [ws])[0]
    ref = O.pdhg_enh(lp, tol=1e-7, x0=first.x, y0=sg * first.row_duals, omega0=first.primal_weight,
                     step=first.step_size)
    got = (int(sol.status), sol.simplex_iterations, sol.restarts, sol.refinements)
    assert got == (ref["status"], ref["iterations"], ref["restarts"], ref["refinements"]), got
    assert sol.status == S.SolutionStatus.Optimal and sol.refinements >= 1
    assert np.array_equal(sol.x, ref["x"]) and np.array_equal(sg * sol.row_duals, ref["y"])
    assert sol.rel_kkt == ref["kkt"] and sol.primal_weight == ref["omega"]
