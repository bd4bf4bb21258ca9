"""FP64 crossover (mpdhg_polish, host code) against the reference's own
SimplexSolver::solve_from_basis and SimplexSolver::solve
(core/src/simplex/simplex_solver.cpp), recorded in
tests/golden/polish_fixtures.{json,npz} by tests/golden/make_polish_fixtures.py.

What pins the polish:
  * status equal to the reference's on every case (optimal, infeasible,
    unbounded);
  * the objective equal to the reference's cold simplex optimum (1e-9
    relative) from every starting basis;
  * an independent FP64 optimality certificate of the returned vertex:
    primal feasibility, dual sign conditions per basis status, nonbasic
    variables at their bounds, and zero duality gap;
  * x equal to the reference's x where the reference's answer is optimal
    and the vertex is unique (every nonbasic reduced cost nonzero).

The reference's polish stops at a primal feasible basis without checking
dual feasibility when it starts its dual simplex from a dual infeasible
basis (simplex_solver.cpp:304-316 has no primal clean-up, unlike solve() at
:168-205), and then reports a non-optimal vertex as optimal.  Three cases
here show it; for them the test asserts, from the reference's own reduced
costs, that its answer is dual infeasible, and that ours reaches the optimum.
CPU only."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import sih_solver as S

G = Path(__file__).parent / "golden"
PF = json.loads((G / "polish_fixtures.json").read_text())
PA = np.load(G / "polish_fixtures.npz")
FA = np.load(G / "fixtures.npz")
CASES = PF["cases"]


def _problem(meta) -> S.Problem:
    a = FA if meta["lp_in"] == "fixtures.npz" else PA
    nm, m, n = meta["lp"], meta["m"], meta["n"]
    p = S.Problem(nm)
    p.resize(m, n)
    p.A = S.SparseMatrix.from_csc(m, n, a[f"{nm}/lp/col_ptr"], a[f"{nm}/lp/row_ind"], a[f"{nm}/lp/vals"])
    p.c, p.col_lower, p.col_upper, p.row_lower, p.row_upper = (
        np.array(a[f"{nm}/lp/{k}"]) for k in ("c", "l", "u", "rl", "ru"))
    return p


def _certificate(p: S.Problem, r: S.PolishResult, tol=1e-9):
    """Optimality of (x, y) at the returned basis, recomputed in numpy."""
    x, y = r.x, r.row_duals
    ax = p.A.mat_vec(x)
    scale = 1.0 + max(np.abs(p.row_lower[np.isfinite(p.row_lower)]).max(initial=0),
                      np.abs(p.row_upper[np.isfinite(p.row_upper)]).max(initial=0),
                      np.abs(x).max(initial=0))
    assert np.all(x >= p.col_lower - tol * scale) and np.all(x <= p.col_upper + tol * scale)
    assert np.all(ax >= p.row_lower - tol * scale) and np.all(ax <= p.row_upper + tol * scale)
    d = p.c - p.A.mat_trans_vec(y)
    assert np.allclose(d, r.reduced_costs, rtol=0, atol=1e-9 * (1 + np.abs(p.c).max()))
    dtol = 1e-7 * (1 + np.abs(p.c).max())
    for j in range(p.num_cols()):
        st = r.col_basis[j]
        if st == S.BasisStatus.Basic:
            assert abs(d[j]) <= dtol
        elif st == S.BasisStatus.AtLower:  # a fixed variable's dual has either sign
            assert x[j] == p.col_lower[j] and (d[j] >= -dtol or p.col_lower[j] == p.col_upper[j])
        elif st == S.BasisStatus.AtUpper:
            assert x[j] == p.col_upper[j] and (d[j] <= dtol or p.col_lower[j] == p.col_upper[j])
        else:
            assert abs(d[j]) <= dtol
    for i in range(p.num_rows()):
        st = r.row_basis[i]
        fixed = p.row_lower[i] == p.row_upper[i]
        if st == S.BasisStatus.AtLower:
            assert (y[i] >= -dtol or fixed) and abs(ax[i] - p.row_lower[i]) <= tol * scale
        elif st == S.BasisStatus.AtUpper:
            assert (y[i] <= dtol or fixed) and abs(ax[i] - p.row_upper[i]) <= tol * scale
        elif st == S.BasisStatus.Basic:
            assert abs(y[i]) <= dtol
    # zero duality gap: c^T x = sum of bound terms of the duals
    dual = 0.0
    for j in range(p.num_cols()):
        if r.col_basis[j] != S.BasisStatus.Basic:
            dual += d[j] * x[j]
    dual += float(y @ ax)
    assert abs(dual - float(p.c @ x)) <= 1e-9 * (1 + abs(float(p.c @ x)))


@pytest.mark.parametrize("meta", CASES, ids=[c["case"] for c in CASES])
def test_polish_from_identified_basis(meta):
    p = _problem(meta)
    k = meta["case"]
    r = S.polish(p, PA[f"{k}/col_basis"], PA[f"{k}/row_basis"])
    assert r.warm_start
    assert int(r.status) == meta["status"] == meta["cold_status"]
    if meta["status"] != 0:
        return
    opt = meta["cold_objective"]
    assert abs(r.objective - opt) <= 1e-9 * max(1.0, abs(opt))
    _certificate(p, r)
    ref_ok = abs(meta["objective"] - opt) <= 1e-9 * max(1.0, abs(opt))
    d = r.reduced_costs
    nonbasic = r.col_basis != S.BasisStatus.Basic
    unique = np.all(np.abs(d[nonbasic]) > 1e-7) and np.all(np.abs(r.row_duals[r.row_basis != 0]) > 1e-7)
    if ref_ok and unique:
        assert np.allclose(r.x, PA[f"{k}/ref/x"], rtol=0, atol=1e-9 * (1 + np.abs(r.x).max()))


def test_reference_polish_defect_is_dual_infeasible():
    """Where the reference's polish reports a worse objective than its own
    cold optimum, its reduced costs violate dual feasibility."""
    bad = [c for c in CASES if c["status"] == 0 and
           c["objective"] > c["cold_objective"] + 1e-9 * max(1.0, abs(c["cold_objective"]))]
    assert len(bad) >= 3
    for c in bad:
        p = _problem(c)
        k = c["case"]
        x = PA[f"{k}/ref/x"]
        dref = PA[f"{k}/ref/reduced_costs"]
        at_lower = np.isclose(x, p.col_lower) & (p.col_lower < p.col_upper)
        at_upper = np.isclose(x, p.col_upper) & (p.col_lower < p.col_upper)
        worst = max((-dref[at_lower]).max(initial=0), dref[at_upper].max(initial=0))
        assert worst > 1e-6, c["case"]
        ours = S.polish(p, PA[f"{k}/col_basis"], PA[f"{k}/row_basis"])
        assert ours.objective < c["objective"]


@pytest.mark.parametrize("lp", sorted({c["lp"] for c in CASES}))
def test_cold_start_reaches_the_reference_optimum(lp):
    meta = next(c for c in CASES if c["lp"] == lp)
    p = _problem(meta)
    r = S.polish(p)
    assert not r.warm_start
    assert int(r.status) == meta["cold_status"]
    if meta["cold_status"] == 0:
        opt = meta["cold_objective"]
        assert abs(r.objective - opt) <= 1e-9 * max(1.0, abs(opt))
        _certificate(p, r)


def test_optimal_basis_takes_no_pivots():
    meta = next(c for c in CASES if c["lp"].startswith("staircase_T4"))
    p = _problem(meta)
    first = S.polish(p)
    again = S.polish(p, first.col_basis, first.row_basis)
    assert again.warm_start and again.iterations == 0 and again.repaired == 0
    assert again.objective == pytest.approx(first.objective, rel=1e-12)


def test_wrong_basis_count_starts_cold():
    meta = CASES[0]
    p = _problem(meta)
    cb = np.zeros(p.num_cols(), np.int8)  # all basic: n != m
    rb = np.zeros(p.num_rows(), np.int8)
    r = S.polish(p, cb, rb)
    assert not r.warm_start and r.status == S.SolutionStatus.Optimal


def test_singular_basis_is_repaired():
    """Two identical basic columns: the factorisation drops one and puts a
    slack in its place, and the solve still ends at the optimum."""
    meta = next(c for c in CASES if c["lp"].startswith("random_m60"))
    p = _problem(meta)
    A = p.A
    m, n = p.num_rows(), p.num_cols()
    cp, ri, v = A.csc_col_ptr(), A.csc_row_ind(), A.csc_values()
    # append a copy of column 0 as a new column n, fixed at 0 so that the LP
    # (and its optimum) does not change
    cp2 = np.concatenate([cp, [cp[-1] + (cp[1] - cp[0])]])
    ri2 = np.concatenate([ri, ri[cp[0]:cp[1]]])
    v2 = np.concatenate([v, v[cp[0]:cp[1]]])
    q = S.Problem("dup")
    q.resize(m, n + 1)
    q.A = S.SparseMatrix.from_csc(m, n + 1, cp2, ri2, v2)
    q.c = np.concatenate([p.c, p.c[:1]])
    q.col_lower = np.concatenate([p.col_lower, [0.0]])
    q.col_upper = np.concatenate([p.col_upper, [0.0]])
    q.row_lower, q.row_upper = p.row_lower, p.row_upper
    cb = np.full(n + 1, S.BasisStatus.AtLower, np.int8)
    rb = np.full(m, S.BasisStatus.Basic, np.int8)
    cb[0] = cb[n] = S.BasisStatus.Basic
    rb[0] = rb[1] = S.BasisStatus.AtLower
    r = S.polish(q, cb, rb)
    assert r.warm_start and r.repaired >= 1
    assert r.status == S.SolutionStatus.Optimal
    assert abs(r.objective - meta["cold_objective"]) <= 1e-9 * max(1.0, abs(meta["cold_objective"]))


def test_iteration_limit_is_reported():
    meta = next(c for c in CASES if c["lp"].startswith("staircase_T4"))
    r = S.polish(_problem(meta), max_iterations=3)
    assert r.status == S.SolutionStatus.IterationLimit and r.iterations == 3


def test_polish_rejects_bad_options():
    p = _problem(CASES[0])
    with pytest.raises(RuntimeError):
        S.polish(p, refactor_every=0)


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
@pytest.mark.parametrize("lp", ["staircase_T4_R40_C90_s6", "staircase_T2_R100_C240_s5",
                                "random_m200_n480_k6_s9"])
def test_crossover_from_a_pdhg_point(lp, tol):
    """The hybrid's crossover: the oracle's PDHG point, its identified basis,
    then the push of every column inside its bounds and the simplex.  The
    objective equals the reference's optimum; from a 1e-8 point, fewer
    pivots follow the push than a cold start takes.  (From a 1e-6 point the
    rows the PDHG leaves 1e-6 off their bounds can cost more pivots than
    that, for the push snaps them onto their bounds.)"""
    meta = next(c for c in CASES if c["lp"] == lp)
    p = _problem(meta)
    e = O.pdhg_enh(O.LpArrays.of(p), tol=tol, max_iterations=400000)
    cb, rb = S.identify_basis(p, e["x"])
    r = S.polish(p, cb, rb, e["x"])
    assert r.status == S.SolutionStatus.Optimal and r.push_steps > 0
    opt = meta["cold_objective"]
    assert abs(r.objective - opt) <= 1e-9 * max(1.0, abs(opt))
    _certificate(p, r)
    if tol <= 1e-8:
        assert r.iterations < S.polish(p).iterations


def test_crossover_start_point_size_is_checked():
    p = _problem(CASES[0])
    with pytest.raises(ValueError):
        S.polish(p, x_start=np.zeros(p.num_cols() + 1))


def _lp(m, n, trip, c, l, u, rl, ru):
    p = S.Problem("edge")
    p.resize(m, n)
    p.A = S.SparseMatrix.from_triplets(m, n, [S.Triplet(*t) for t in trip])
    p.c, p.col_lower, p.col_upper = (np.array(v, float) for v in (c, l, u))
    p.row_lower, p.row_upper = np.array(rl, float), np.array(ru, float)
    return p


def test_polish_without_rows():
    """m = 0: every column sits at the bound its cost prefers."""
    p = _lp(0, 3, [], [1.0, -2.0, 0.0], [0.0, -1.0, 2.0], [5.0, 4.0, 7.0], [], [])
    r = S.polish(p)
    assert r.status == S.SolutionStatus.Optimal
    assert r.x.tolist() == [0.0, 4.0, 2.0] and r.objective == -8.0


def test_polish_unbounded_and_infeasible_edges():
    # min -x, x >= 0 free above, one row x - y = 0 with y >= 0: unbounded
    p = _lp(1, 2, [(0, 0, 1.0), (0, 1, -1.0)], [-1.0, 0.0], [0.0, 0.0], [np.inf, np.inf], [0.0], [0.0])
    assert S.polish(p).status == S.SolutionStatus.Unbounded
    # x + y = 5 with x, y in [0, 1]: infeasible
    q = _lp(1, 2, [(0, 0, 1.0), (0, 1, 1.0)], [1.0, 1.0], [0.0, 0.0], [1.0, 1.0], [5.0], [5.0])
    assert S.polish(q).status == S.SolutionStatus.Infeasible


def test_polish_free_and_fixed_columns():
    # min x0 + x1 s.t. x0 - x1 = 2, x0 free, x1 fixed at 3  ->  x0 = 5
    p = _lp(1, 2, [(0, 0, 1.0), (0, 1, -1.0)], [1.0, 1.0], [-np.inf, 3.0], [np.inf, 3.0], [2.0], [2.0])
    r = S.polish(p)
    assert r.status == S.SolutionStatus.Optimal
    assert np.allclose(r.x, [5.0, 3.0]) and abs(r.objective - 8.0) < 1e-12
