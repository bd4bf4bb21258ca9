"""Writes tests/golden/fixtures.npz + fixtures.json from the reference's own
FP64 dual simplex (core/src/simplex, built by oracle/ref/Makefile into
oracle/_ref/refsimplex).  Run in the build container (it needs the
reference sources to build the solver):

    make -C oracle/ref && python tests/golden/make_fixtures.py

Fixture LPs: the LP toys the reference ships (data/toy/toy01, toy04, toy05,
read by the reference's MPS reader) and small synthetic LPs from this
engine's generators.  Stored per LP: the LP data, and the simplex status,
objective, x, row duals (simplex sign: reduced cost = c - A^T y), reduced
costs and basis.
"""
from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "refinery-optimizer_amd"))
import sih_solver as S  # noqa: E402

REFSIMPLEX = ROOT / "oracle" / "_ref" / "refsimplex"
TOYS = Path("/root/reference/data/toy")


def parse_solution(text: str) -> dict:
    out = {}
    for line in text.strip().splitlines():
        key, *vals = line.split()
        if key in ("status", "iterations"):
            out[key] = int(vals[0])
        elif key == "objective":
            out[key] = float(vals[0])
        elif key in ("col_basis", "row_basis"):
            out[key] = np.array([int(v) for v in vals], np.int8)
        else:
            out[key] = np.array([float(v) for v in vals], np.float64)
    return out


def parse_dump(path: Path) -> dict:
    toks = path.read_text().split()
    pos = 0

    def take(k, cast):
        nonlocal pos
        vals = [cast(t) for t in toks[pos:pos + k]]
        pos += k
        return vals

    assert toks[0] == "sense"
    sense = int(toks[1])
    pos = 2
    m, n, nnz = take(3, int)
    f = lambda t: float(t)  # noqa: E731  ("inf" parses)
    lp = dict(m=m, n=n, col_ptr=np.array(take(n + 1, int), np.int64),
              row_ind=np.array(take(nnz, int), np.int64), vals=np.array(take(nnz, f)),
              c=np.array(take(n, f)), l=np.array(take(n, f)), u=np.array(take(n, f)),
              rl=np.array(take(m, f)), ru=np.array(take(m, f)))
    return sense, lp


def write_lp(path: Path, p: S.Problem) -> None:
    def fmt(a):
        return " ".join("inf" if v == np.inf else "-inf" if v == -np.inf else repr(float(v)) for v in a)

    A = p.A
    path.write_text("\n".join([
        f"{p.num_rows()} {p.num_cols()} {p.num_nonzeros()}",
        " ".join(map(str, A.csc_col_ptr())), " ".join(map(str, A.csc_row_ind())),
        fmt(A.csc_values()), fmt(p.c), fmt(p.col_lower), fmt(p.col_upper), fmt(p.row_lower),
        fmt(p.row_upper)]) + "\n")


def main():
    arrays, meta = {}, []
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        cases = []
        for toy in ("toy01_lp_simple", "toy04_lp_equality_ranges", "toy05_lp_bounds"):
            dump = td / f"{toy}.txt"
            res = subprocess.run([str(REFSIMPLEX), "--mps", str(TOYS / f"{toy}.mps"), str(dump)],
                                 check=True, capture_output=True, text=True)
            sense, lp = parse_dump(dump)
            cases.append((toy, f"reference data/toy/{toy}.mps", sense, lp, parse_solution(res.stdout)))
        # The reference LU recomputes column degrees per pivot (O(m^3) per
        # factorisation, sparse_lu.cpp:238-251): keep the synthetic cases small.
        gens = [("random_m60_n150_k4_s0", lambda: S.gen_random(60, 150, 4, 0)),
                ("random_m120_n288_k8_s1", lambda: S.gen_random(120, 288, 8, 1)),
                ("staircase_T1_R40_C96_s2", lambda: S.gen_staircase(1, 40, 96, 2)),
                ("staircase_T3_R40_C90_s4", lambda: S.gen_staircase(3, 40, 90, 4)),
                ("staircase_T2_R100_C240_s5", lambda: S.gen_staircase(2, 100, 240, 5))]
        for name, fn in gens:
            p = fn()
            path = td / f"{name}.txt"
            write_lp(path, p)
            res = subprocess.run([str(REFSIMPLEX), str(path)], check=True, capture_output=True,
                                 text=True, timeout=300)
            A = p.A
            lp = dict(m=p.num_rows(), n=p.num_cols(), col_ptr=A.csc_col_ptr(),
                      row_ind=A.csc_row_ind(), vals=A.csc_values(), c=p.c, l=p.col_lower,
                      u=p.col_upper, rl=p.row_lower, ru=p.row_upper)
            cases.append((name, "sih_solver generator", 1, lp, parse_solution(res.stdout)))
        for name, source, sense, lp, sol in cases:
            for k in ("col_ptr", "row_ind", "vals", "c", "l", "u", "rl", "ru"):
                arrays[f"{name}/lp/{k}"] = np.asarray(lp[k])
            for k in ("x", "row_duals", "reduced_costs", "col_basis", "row_basis"):
                arrays[f"{name}/ref/{k}"] = sol[k]
            meta.append(dict(name=name, source=source, sense=sense, m=lp["m"], n=lp["n"],
                             status=sol["status"], objective=sol["objective"],
                             iterations=sol["iterations"]))
            print(f"{name}: status {sol['status']} objective {sol['objective']:.12g}")
    out = Path(__file__).resolve().parent
    np.savez_compressed(out / "fixtures.npz", **arrays)
    (out / "fixtures.json").write_text(json.dumps(
        {"solver": "reference FP64 dual simplex (core/src/simplex) via oracle/ref/ref_driver.cpp",
         "dual_sign": "simplex: reduced_cost = c - A^T row_duals", "cases": meta}, indent=1) + "\n")


if __name__ == "__main__":
    main()
