"""Fixtures written by the reference's own FP64 dual simplex
(tests/golden/make_fixtures.py) pin the oracle, the KKT evaluator and the
engine's host-side basis identification."""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

import oracle as O
import sih_solver as S

GOLD = Path(__file__).resolve().parent / "golden"
META = json.loads((GOLD / "fixtures.json").read_text())["cases"]
ARR = np.load(GOLD / "fixtures.npz", allow_pickle=False)
NAMES = [c["name"] for c in META]


def fixture(name):
    meta = next(c for c in META if c["name"] == name)
    g = lambda k: ARR[f"{name}/lp/{k}"]  # noqa: E731
    lp = O.LpArrays(meta["m"], meta["n"], g("col_ptr"), g("row_ind"), g("vals"), g("c"), g("l"),
                    g("u"), g("rl"), g("ru"))
    ref = {k: ARR[f"{name}/ref/{k}"] for k in ("x", "row_duals", "reduced_costs", "col_basis",
                                                "row_basis")}
    return meta, lp, ref


def problem_of(lp: O.LpArrays) -> S.Problem:
    p = S.Problem("fixture")
    p.resize(lp.m, lp.n)
    p.A = S.SparseMatrix.from_csc(lp.m, lp.n, lp.cp, lp.ri, lp.v)
    p.c, p.col_lower, p.col_upper, p.row_lower, p.row_upper = lp.c, lp.l, lp.u, lp.rl, lp.ru
    return p


def test_reference_metadata_toys():
    # data/toy/toy_suite_metadata.json:5-12 (toy01) and the toy04 entry
    m1, _, r1 = fixture("toy01_lp_simple")
    assert m1["objective"] == -6.5 and np.allclose(r1["x"], [1.5, 2.5], atol=1e-12)
    m4, _, r4 = fixture("toy04_lp_equality_ranges")
    assert m4["objective"] == 16.0 and np.allclose(r4["x"], [0.0, 4.0], atol=1e-12)


@pytest.mark.parametrize("name", NAMES)
def test_simplex_point_is_kkt_optimal_in_pdhg_sign(name):
    meta, lp, ref = fixture(name)
    y_pdhg = -ref["row_duals"]  # simplex: c - A^T y ; PDHG: c + A^T y
    k = O.kkt(lp, ref["x"], y_pdhg)
    assert k["kkt"] < 1e-9, k
    assert abs(k["pobj"] - meta["objective"]) <= 1e-9 * (1 + abs(meta["objective"]))
    lam = lp.c + sp.csc_matrix((lp.v, lp.ri, lp.cp), shape=(lp.m, lp.n)).T @ y_pdhg
    np.testing.assert_allclose(lam, ref["reduced_costs"], atol=1e-9)


@pytest.mark.parametrize("name", NAMES)
def test_highs_agrees_with_reference_simplex(name):
    meta, lp, _ = fixture(name)
    A = sp.csc_matrix((lp.v, lp.ri, lp.cp), shape=(lp.m, lp.n))
    lo = np.where(np.isfinite(lp.rl), lp.rl, -np.inf)
    hi = np.where(np.isfinite(lp.ru), lp.ru, np.inf)
    from scipy.optimize import LinearConstraint, milp, Bounds

    res = milp(lp.c, constraints=[LinearConstraint(A, lo, hi)], bounds=Bounds(lp.l, lp.u))
    assert res.status == 0
    assert abs(res.fun - meta["objective"]) <= 1e-8 * (1 + abs(meta["objective"]))


@pytest.mark.parametrize("name", NAMES)
def test_enhanced_port_reaches_simplex_objective(name):
    meta, lp, _ = fixture(name)
    r = O.pdhg_enh(lp, tol=1e-8, max_iterations=400000, max_refinements=8)
    assert r["status"] == 0 and r["kkt"] <= 1e-8
    # the objective error can exceed the KKT tolerance by |y| * |Ax - b|: allow 100x
    assert abs(r["primal_obj"] - meta["objective"]) <= 1e-6 * (1 + abs(meta["objective"]))


@pytest.mark.parametrize("name", ["toy01_lp_simple", "toy04_lp_equality_ranges", "random_m60_n150_k4_s0",
                                  "staircase_T1_R40_C96_s2"])
def test_reference_pdhg_restatement_fixed_point(name):
    """The reference-mode restatement, run long in FP64, stops only near the
    simplex optimum (its own stopping rule is the reference's P and D)."""
    meta, lp, _ = fixture(name)
    r = O.pdhg_ref(lp, max_iterations=300000, tol_primal=1e-7, tol_dual=1e-7, fp64=True)
    if r["status"] != 0:
        pytest.skip(f"reference method did not stop within 300k iterations ({r['last_primal']:.1e}, "
                    f"{r['last_dual']:.1e})")
    assert abs(r["objective"] - meta["objective"]) <= 1e-4 * (1 + abs(meta["objective"]))


@pytest.mark.parametrize("name", NAMES)
def test_basis_identification_recovers_simplex_basis(name):
    """At a nondegenerate optimal vertex the m most interior variables are
    exactly the simplex's basic ones (gpu_solver.cpp:334-395)."""
    meta, lp, ref = fixture(name)
    x = ref["x"]
    A = sp.csc_matrix((lp.v, lp.ri, lp.cp), shape=(lp.m, lp.n))
    ax = A @ x
    dist = np.concatenate([np.minimum(np.where(np.isfinite(lp.l), x - lp.l, np.inf),
                                      np.where(np.isfinite(lp.u), lp.u - x, np.inf)),
                           np.minimum(np.where(np.isfinite(lp.rl), ax - lp.rl, np.inf),
                                      np.where(np.isfinite(lp.ru), lp.ru - ax, np.inf))])
    basic = np.concatenate([ref["col_basis"], ref["row_basis"]]) == 0
    if (dist[basic] <= 1e-9).any() or (dist[~basic] > 1e-9).any():
        pytest.skip("degenerate vertex: the ranking is not unique")
    cb, rb = O.identify_basis(lp, x)
    got = np.concatenate([cb, rb]) == 0
    assert np.array_equal(got, basic)
    pcb, prb = S.identify_basis(problem_of(lp), x)  # engine host code, no GPU needed
    assert np.array_equal(pcb, cb) and np.array_equal(prb, rb)


def test_engine_basis_identification_equals_oracle_on_interior_points():
    rng = np.random.default_rng(5)
    for name in NAMES:
        _, lp, _ = fixture(name)
        for _ in range(3):
            lo = np.where(np.isfinite(lp.l), lp.l, -5.0)
            hi = np.where(np.isfinite(lp.u), lp.u, 5.0)
            x = lo + (hi - lo) * rng.random(lp.n)
            at = rng.random(lp.n) < 0.3  # put some columns on a bound: ties at distance 0
            x[at] = lo[at]
            cb, rb = O.identify_basis(lp, x)
            pcb, prb = S.identify_basis(problem_of(lp), x)
            assert np.array_equal(pcb, cb) and np.array_equal(prb, rb), name
