"""GpuSolver::solve_hybrid end to end (gpu_solver.cpp:309-417): PDHG on the
GPU, the basis hand-off, the FP64 crossover on the host.

Checked against the reference's own results: the toy smoke test of
tests/test_gpu_basic.py:33-58 (hybrid optimal, objective equal to the CPU
simplex's), the statuses and optima of the reference's simplex on the polish
fixtures (tests/golden/polish_fixtures.json), and the HiGHS optimum of the
12-block staircase (tests/golden/highs_staircase.json)."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import sih_solver as S

pytestmark = pytest.mark.gpu

G = Path(__file__).parent / "golden"
MPS = {c["name"]: c for c in json.loads((G / "mps_fixtures.json").read_text())["cases"]}
BASE = {c["name"]: c for c in json.loads((G / "fixtures.json").read_text())["cases"]}
PF = json.loads((G / "polish_fixtures.json").read_text())["cases"]
PA = np.load(G / "polish_fixtures.npz")
FA = np.load(G / "fixtures.npz")


def _toy(tmp_path, name) -> S.Problem:
    path = tmp_path / f"{name}.mps"
    path.write_text(MPS[name]["text"])
    return S.read_mps(str(path))


def _fixture_lp(meta) -> S.Problem:
    a = FA if meta["lp_in"] == "fixtures.npz" else PA
    nm, m, n = meta["lp"], meta["m"], meta["n"]
    p = S.Problem(nm)
    p.resize(m, n)
    p.A = S.SparseMatrix.from_csc(m, n, a[f"{nm}/lp/col_ptr"], a[f"{nm}/lp/row_ind"], a[f"{nm}/lp/vals"])
    p.c, p.col_lower, p.col_upper, p.row_lower, p.row_upper = (
        np.array(a[f"{nm}/lp/{k}"]) for k in ("c", "l", "u", "rl", "ru"))
    return p


def test_hybrid_toy_smoke_as_the_reference_runs_it(tmp_path):
    """tests/test_gpu_basic.py:33-58: toy01 read from MPS, 2000 iterations at
    1e-4, the hybrid optimal and within 1e-3 of the CPU simplex (here 1e-9)."""
    prob = _toy(tmp_path, "toy01_lp_simple")
    cfg = S.GpuPdhgConfig()
    cfg.max_iterations = 2000
    cfg.tol_primal = 1e-4
    sol_gpu = S.solve_gpu_pdhg(prob, cfg)
    assert np.all(np.isfinite(sol_gpu.x))
    sol_hybrid = S.solve_gpu_hybrid(prob, cfg)
    assert sol_hybrid.is_optimal()
    assert abs(sol_hybrid.primal_objective - BASE["toy01_lp_simple"]["objective"]) < 1e-9
    ref_x = FA["toy01_lp_simple/ref/x"]
    assert np.allclose(sol_hybrid.x, ref_x, atol=1e-9)
    assert len(sol_hybrid.col_basis) == prob.num_cols()


@pytest.mark.parametrize("toy,status", [("toy02_lp_infeasible", S.SolutionStatus.Infeasible),
                                        ("toy03_lp_unbounded", S.SolutionStatus.Unbounded)])
def test_hybrid_reports_the_reference_status(tmp_path, toy, status):
    prob = _toy(tmp_path, toy)
    sol = S.solve_gpu_hybrid(prob, S.GpuPdhgConfig(max_iterations=500))
    assert sol.status == status


LPS = sorted({c["lp"]: c for c in PF}.values(), key=lambda c: c["lp"])


@pytest.mark.parametrize("mode", ["reference", "enhanced"])
@pytest.mark.parametrize("meta", [c for c in LPS if c["cold_status"] == 0], ids=lambda c: c["lp"])
def test_hybrid_reaches_the_reference_optimum(meta, mode):
    p = _fixture_lp(meta)
    cfg = S.GpuPdhgConfig() if mode == "reference" else S.GpuPdhgConfig.enhanced(tol=1e-4)
    sol = S.solve_gpu_hybrid(p, cfg)
    opt = meta["cold_objective"]
    assert sol.is_optimal()
    assert abs(sol.primal_objective - opt) <= 1e-9 * max(1.0, abs(opt))
    ax = p.A.mat_vec(sol.x)
    assert np.all(ax >= p.row_lower - 1e-7 * (1 + np.abs(ax))) and np.all(ax <= p.row_upper + 1e-7 * (1 + np.abs(ax)))


def test_hybrid_staircase12_matches_highs():
    """12-block staircase (4800 x 10800): enhanced PDHG to 1e-6, then the
    crossover; the objective equals HiGHS's to 1e-9 and few pivots remain
    after the push."""
    ref = next(r for r in json.loads((G / "highs_staircase.json").read_text()) if r["T"] == 12)
    p = S.gen_staircase(ref["T"], ref["R"], ref["C"], ref["seed"])
    sol = S.solve_gpu_hybrid(p, S.GpuPdhgConfig.enhanced(tol=1e-6))
    assert sol.is_optimal()
    assert abs(sol.primal_objective - ref["objective"]) <= 1e-9 * abs(ref["objective"])
    assert sol.simplex_iterations < 2000
