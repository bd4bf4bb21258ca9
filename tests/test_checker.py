"""mpdhg_check_solution against the reference's own solution checker.

tests/golden/checker_fixtures.json holds what checker::check_solution
(core/src/checker/checker.cpp:9-268, built from the reference sources) reported
on crafted candidates (make_checker_fixtures.py): every field must be equal
-- doubles bit for bit, flags, and the summary text -- since the library
restates the checker with the same arithmetic order.  A further check shows
the checker accepts the engine's own kind of result: the reference-simplex
optimum of an LP passes, and the same point with the PDHG dual sign fails
stationarity until the duals are negated.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import sih_solver as S

FIX = json.loads((Path(__file__).parent / "golden" / "checker_fixtures.json").read_text())
INTS = ("is_primal_feasible", "is_dual_feasible", "is_complementary", "is_integrality_satisfied",
        "is_certificate_valid", "all_checks_passed")


def _problem(case):
    lp = case["lp"]
    p = S.Problem(case["name"])
    p.resize(lp["m"], lp["n"])
    p.A = S.SparseMatrix.from_csc(lp["m"], lp["n"], np.array(lp["col_ptr"], np.int64),
                                  np.array(lp["row_ind"], np.int64), np.array(lp["vals"], np.float64))
    p.c = np.array(lp["c"], float)
    p.col_lower, p.col_upper = np.array(lp["col_lower"], float), np.array(lp["col_upper"], float)
    p.row_lower, p.row_upper = np.array(lp["row_lower"], float), np.array(lp["row_upper"], float)
    p.sense = S.ObjectiveSense(case["sense"])
    p.obj_offset = case["offset"]
    p.var_types = np.array(case["vtypes"], np.int8)
    if case["q"] is not None:
        cp, ri, qv = case["q"]
        p.Q = S.SparseMatrix.from_csc(lp["n"], lp["n"], np.array(cp, np.int64), np.array(ri, np.int64),
                                      np.array(qv, float))
    return p


def _solution(case):
    sol = case["solution"]
    s = S.Solution()
    s.status = S.SolutionStatus(sol["status"])
    s.primal_objective = sol["objective"]
    for k in ("x", "row_duals", "reduced_costs"):
        setattr(s, k, np.array(sol.get(k, []), float))
    return s, sol.get("ray")


@pytest.mark.parametrize("case", FIX, ids=[c["name"] for c in FIX])
def test_checker_matches_the_reference_checker(case):
    p = _problem(case)
    s, ray = _solution(case)
    r = S.check_solution(p, s, *case["tols"], ray=ray)
    exp = case["expected"]
    for k, v in exp.items():
        got = getattr(r, k)
        if k == "summary":
            assert got == v, k
        elif k in INTS:
            assert int(got) == v, k
        else:
            e = float.fromhex(v)
            assert (math.isnan(e) and math.isnan(got)) or got == e, (k, got, e)


def test_checker_reads_the_simplex_dual_sign():
    case = next(c for c in FIX if c["name"] == "random_lp_optimal")
    p = _problem(case)
    s, _ = _solution(case)
    assert S.check_solution(p, s).is_dual_feasible
    s.row_duals = -s.row_duals  # the PDHG sign: reduced cost = c + A^T y
    r = S.check_solution(p, s)
    assert not r.is_dual_feasible and r.max_dual_residual > 1e-3


def test_checker_rejects_malformed_input():
    p = S.gen_random(20, 50, 4, 1)
    s = S.Solution(x=np.zeros(50))
    p.A._ri = p.A._ri.copy()
    p.A._ri[3] = 99  # row index out of range
    with pytest.raises(RuntimeError, match="row index out of range"):
        S.check_solution(p, s)
