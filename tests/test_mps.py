"""MPS / QPS reader and writer (libmpdhg, host code) against the reference's
own io::read_mps / io::write_mps (core/src/io/mps_io.cpp), whose outputs on
the reference's data/toy files and on crafted inputs are stored in
tests/golden/mps_fixtures.json (tests/golden/make_mps_fixtures.py).

The bar is exact: every parsed field bit for bit, every written file byte for
byte.  CPU only."""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

import sih_solver as S

FIX = json.loads((Path(__file__).parent / "golden" / "mps_fixtures.json").read_text())
CASES = {c["name"]: c for c in FIX["cases"]}
PARSED = [c for c in FIX["cases"] if "parse" in c]
ERRORS = [c for c in FIX["cases"] if "error" in c]
GENERATED = [c for c in FIX["cases"] if "parse" not in c and "error" not in c]


def _f(tok: str) -> float:
    if tok in ("inf", "-inf", "nan"):
        return float(tok)
    return float.fromhex(tok)


def _same(a, b) -> bool:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.shape != b.shape:
        return False
    # bit-exact, NaN equal to NaN, and -0.0 distinct from 0.0
    return bool(np.all(a.view(np.uint64) == b.view(np.uint64)) or
                all((math.isnan(x) and math.isnan(y)) or (x == y and math.copysign(1, x) == math.copysign(1, y))
                    for x, y in zip(a.tolist(), b.tolist())))


def _read_text(tmp_path, case) -> S.Problem:
    path = tmp_path / f"{case['name']}.mps"
    path.write_bytes(case["text"].encode())
    return S.read_mps(str(path))


def test_fixture_inventory():
    names = set(CASES)
    for toy in ("toy01_lp_simple", "toy02_lp_infeasible", "toy03_lp_unbounded",
                "toy04_lp_equality_ranges", "toy05_lp_bounds", "toy06_milp_knapsack",
                "toy07_milp_mixed_integer", "toy08_qp_convex_diagonal", "toy09_qp_convex_general",
                "toy10_milp_infeasible_integer"):
        assert toy in names
    assert len(ERRORS) >= 4 and len(GENERATED) >= 2


@pytest.mark.parametrize("case", PARSED, ids=[c["name"] for c in PARSED])
def test_read_matches_reference_reader(tmp_path, case):
    ref = case["parse"]
    p = _read_text(tmp_path, case)
    assert p.name == ref["name"]
    assert int(p.sense) == ref["sense"]
    assert _same([p.obj_offset], [_f(ref["obj_offset"])])
    assert (p.num_rows(), p.num_cols(), p.num_nonzeros()) == (ref["m"], ref["n"], ref["nnz"])
    A = p.A
    assert A.csc_col_ptr().tolist() == ref["col_ptr"]
    assert A.csc_row_ind().tolist() == ref["row_ind"]
    assert _same(A.csc_values(), [_f(t) for t in ref["vals"]])
    for ours, key in ((p.c, "c"), (p.col_lower, "col_lower"), (p.col_upper, "col_upper"),
                      (p.row_lower, "row_lower"), (p.row_upper, "row_upper")):
        assert _same(ours, [_f(t) for t in ref[key]]), key
    assert p.var_types.tolist() == ref["var_types"]
    assert p.row_names == ref["row_names"]
    assert p.col_names == ref["col_names"]
    assert p.has_quadratic() == bool(ref["has_quadratic"])
    if ref["has_quadratic"]:
        assert p.Q.csc_col_ptr().tolist() == ref["q_col_ptr"]
        assert p.Q.csc_row_ind().tolist() == ref["q_row_ind"]
        assert _same(p.Q.csc_values(), [_f(t) for t in ref["q_vals"]])


@pytest.mark.parametrize("case", PARSED, ids=[c["name"] for c in PARSED])
def test_write_matches_reference_writer(tmp_path, case):
    p = _read_text(tmp_path, case)
    out = tmp_path / "out.mps"
    S.write_mps(p, str(out))
    assert out.read_text() == case["written"]


@pytest.mark.parametrize("case", ERRORS, ids=[c["name"] for c in ERRORS])
def test_read_errors_match_reference(tmp_path, case):
    with pytest.raises(RuntimeError) as ei:
        _read_text(tmp_path, case)
    msg = str(ei.value)
    # the reference's text up to its first detail (our message stops before
    # the ": lower=... > upper=..." suffix of Problem::validate)
    assert case["error"].split(":")[0] in msg


def test_missing_file(tmp_path):
    with pytest.raises(RuntimeError, match="Could not open MPS file"):
        S.read_mps(str(tmp_path / "absent.mps"))


@pytest.mark.parametrize("case", GENERATED, ids=[c["name"] for c in GENERATED])
def test_writer_output_is_the_reference_writers_fixed_point(tmp_path, case):
    """Generated LPs written here, read by the reference reader and written
    again by the reference writer give the same text."""
    parts = case["name"].split("_")
    if parts[1] == "staircase":
        T, R, Cb, s = (int(t[1:]) for t in parts[2:6])
        p = S.gen_staircase(T, R, Cb, s)
    else:
        m, n, k, s = (int(t[1:]) for t in parts[2:6])
        p = S.gen_random(m, n, k, s)
    out = tmp_path / "gen.mps"
    S.write_mps(p, str(out))
    assert out.read_text() == case["written"]


def test_round_trip_keeps_the_lp(tmp_path):
    """write -> read reproduces an LP whose values print exactly in 12 digits."""
    p = S.Problem("rt")
    p.resize(3, 4)
    trip = [S.Triplet(0, 0, 1.5), S.Triplet(1, 0, -2.0), S.Triplet(2, 1, 0.25), S.Triplet(0, 2, 3.0),
            S.Triplet(1, 3, 7.125), S.Triplet(2, 3, -1.0)]
    p.A = S.SparseMatrix.from_triplets(3, 4, trip)
    p.c = np.array([1.0, -2.0, 0.0, 4.5])
    p.row_lower = np.array([-np.inf, 1.0, 2.0])
    p.row_upper = np.array([5.0, 1.0, 6.5])
    p.col_lower = np.array([0.0, -np.inf, -3.0, 1.0])
    p.col_upper = np.array([np.inf, np.inf, 8.0, 1.0])
    p.row_names = ["a", "b", "c"]
    p.col_names = ["w", "x", "y", "z"]
    p.obj_offset = 2.5
    path = tmp_path / "rt.mps"
    S.write_mps(p, str(path))
    q = S.read_mps(str(path))
    assert q.name == "rt" and q.row_names == p.row_names and q.col_names == p.col_names
    assert q.obj_offset == 2.5
    for a, b in ((q.c, p.c), (q.row_lower, p.row_lower), (q.row_upper, p.row_upper),
                 (q.col_lower, p.col_lower), (q.col_upper, p.col_upper),
                 (q.A.csc_values(), p.A.csc_values())):
        assert np.array_equal(a, b)
    assert np.array_equal(q.A.csc_row_ind(), p.A.csc_row_ind())


def test_mps_lp_feeds_the_solver_view(tmp_path):
    """A parsed toy produces the C view the solver reads (no device needed)."""
    q = _read_text(tmp_path, CASES["toy01_lp_simple"])
    lp, keep = q._c_view()
    assert (lp.m, lp.n, lp.nnz) == (2, 2, 4)


def test_empty_file_reads_as_an_empty_problem(tmp_path):
    path = tmp_path / "empty.mps"
    path.write_text("")
    p = S.read_mps(str(path))
    assert (p.num_rows(), p.num_cols(), p.num_nonzeros()) == (0, 0, 0)
    assert p.name == "problem" and p.sense == S.ObjectiveSense.Minimize
