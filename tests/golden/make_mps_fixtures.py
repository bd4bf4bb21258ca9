"""Writes tests/golden/mps_fixtures.json from the reference's own MPS reader
and writer (core/src/io/mps_io.cpp, built by oracle/ref/Makefile into
oracle/_ref/refsimplex).  Run in the build container:

    make -C oracle/ref && python tests/golden/make_mps_fixtures.py

Inputs: the MPS files the reference ships under data/toy (their text is kept
in the fixture as input data), and small MPS texts written here to reach
each rule of the reader (ranges, markers, bound types, duplicates, quadratic
sections, comments, CRLF, unknown headers, errors).  Stored per input: the
reference reader's parse of every model::Problem field (floats as C99 hex),
or its error text; and the text the reference writer writes for that parse.
A last group stores the reference writer's text for LPs that this library's
generators make, after a trip through this library's writer and the
reference reader (the writer's output is a fixed point when both writers
agree).
"""
from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "refinery-optimizer_amd"))
import sih_solver as S  # noqa: E402

REFSIMPLEX = ROOT / "oracle" / "_ref" / "refsimplex"
TOYS = Path("/root/reference/data/toy")

CRAFTED = {
    "ranges_rhs": """NAME          RANGES_RHS
ROWS
 N  COST
 G  G1
 G  G2
 L  L1
 L  L2
 E  E1
 E  E2
 E  E3
 L  NORHS
COLUMNS
    X         COST      1.5        G1        1.0
    X         G2        2.0        L1        -1.0
    Y         L2        3.0        E1        1.0
    Y         E2        1.0        E3        0.5
    Y         NORHS     1.0        COST      -2.25
RHS
    RHS       G1        1.0        G2        2.0
    RHS       L1        3.0        L2        4.0
    E1        5.0
    RHS       E2        6.0        E3        7.0
    RHS       COST      12.5
RANGES
    RNG       G1        2.0        G2        -3.0
    RNG       L1        4.0        L2        -5.0
    RNG       E2        8.0
    E3        -9.0
    RNG       E1        -0.0
ENDATA
""",
    "markers_bounds": """NAME          MARKERS
ROWS
 N  OBJ
 L  R1
COLUMNS
    A         OBJ       1.0        R1        1.0
    MARKER                 'MARKER'                 'INTORG'
    B         OBJ       2.0        R1        1.0
    C         R1        1.0
    MARKER                 'MARKER'                 'INTEND'
    D         OBJ       -1.0       R1        1.0
    M2        'MARKER'  'intorg'
    E         R1        1.0
    M3        'MARKER'  'intend'
    F         R1        2.0
    G         R1        1.0
    H         R1        1.0
    I         R1        1.0
    J         R1        1.0
RHS
    RHS       R1        10.0
BOUNDS
 UP BND       A         4.0
 LO BND       B         -1.0
 UI BND       C         7.0
 MI BND       D
 FX BND       E         3.0
 FR BND       F
 BV BND       G
 LI BND       H         2.0
 PL BND       I
 UP BND       J         -2.0
 LO BND       J         -5.0
 up bnd       NEWCOL    6.0
 SC BND       OTHER     1.0
 UP BND       K0
ENDATA
""",
    "duplicates_zeros": """NAME          DUPS
ROWS
 N  OBJ
 L  R1
 G  R2
 N  OBJ2
 E  R1
 L  R3
COLUMNS
    X1        OBJ       1.0        R1        1.0
    X1        R1        2.0        R2        1.0
    X1        OBJ       0.5        OBJ2      9.0
    X2        R2        1e-16      R3        1.0
    X2        R1        0.25       R1        -0.25
    X2        NOROW     4.0
    X1        R3        3.0
    X3        R2        1.0
RHS
    RHS       R1        4.0        R2        -1.0
    RHS       R3        2.0        NOROW     1.0
ENDATA
""",
    "quadratic": """NAME          QUAD
ROWS
 N  OBJ
 L  R1
COLUMNS
    X1        OBJ       -1.0       R1        1.0
    X2        OBJ       -2.0       R1        1.0
RHS
    RHS       R1        4.0
QUADOBJ
    X1        X1        2.0
    X1        X2        0.5
    X2        X2        1.0
    X3        X3        3.0
QMATRIX
    X1        X2        0.25
    X2        X1        0.25
ENDATA
""",
    "format_crlf": "NAME          FMT\r\n* a comment line\r\nrows\r\n N  OBJ\r\n\tL  R1\r\n  G  'ROW 2'\r\n\r\n"
                   "columns\r\n    X1\tOBJ\t1.0\tR1\t1.0\r\n    X2  'ROW 2'  2.0  OBJ  -3.0\r\n"
                   "FOOBAR\r\n    X3  R1  1.0\r\nrhs\r\n    RHS  R1  5  'ROW 2'  1.5e0\r\n"
                   "bounds\r\n UP BND X1 2.5\r\nendata\r\n    X9  R1  1.0\r\n",
    "objsense_max": """NAME
OBJSENSE
    MAXIMIZE
ROWS
 N  OBJ
 L  C1
COLUMNS
    X         OBJ       1.0        C1        1.0
RHS
    RHS       C1        1.0
ENDATA
""",
    "objsense_inline": """NAME          INLINE extra tokens
OBJSENSE MAX
ROWS
 N  OBJ
 G  C1
COLUMNS
    X         OBJ       1.0        C1        1.0
RHS
    RHS       C1        1.0
ENDATA
""",
    "bounds_only_columns": """NAME          NOCOLS
ROWS
 N  OBJ
 L  R1
 G  R2
 E  R3
 X  FREE1
RHS
    RHS       R1        2.0        R2        -3.0
    RHS       R3        0.5        FREE1     1.0
RANGES
    RNG       R2        1.0
BOUNDS
 UP BND       V1        4.0
 FR BND       V2
ENDATA
""",
    "no_endata": """NAME          NOEND
ROWS
 N  OBJ
 E  R1
COLUMNS
    X         OBJ       1e300      R1        -1e-300
    Y         R1        0x1.8p1
RHS
    RHS       R1        inf
""",
    "long_names_int_bounds": """NAME          LONG_NAMES_AND_INTEGER_BOUNDS
ROWS
 N  OBJECTIVE_ROW
 L  CAPACITY_CONSTRAINT_1
 G  DEMAND_CONSTRAINT_22
COLUMNS
    MARKER                 'MARKER'                 'INTORG'
    PRODUCTION_LEVEL_A  OBJECTIVE_ROW  2.5   CAPACITY_CONSTRAINT_1  1.0
    PRODUCTION_LEVEL_A  DEMAND_CONSTRAINT_22  1.0
    MARKER                 'MARKER'                 'INTEND'
    INVENTORY_HOLDING_B  OBJECTIVE_ROW  0.125  DEMAND_CONSTRAINT_22  -1.0
    X  CAPACITY_CONSTRAINT_1  1e-3
RHS
    RHS  CAPACITY_CONSTRAINT_1  1234.5678901234  DEMAND_CONSTRAINT_22  -7.25e-5
BOUNDS
 LO BND  PRODUCTION_LEVEL_A  -3
 UP BND  PRODUCTION_LEVEL_A  40
 MI BND  INVENTORY_HOLDING_B
 UP BND  INVENTORY_HOLDING_B  12.000000000001
ENDATA
""",
    "error_number": """NAME          BADNUM
ROWS
 N  OBJ
 L  R1
COLUMNS
    X         OBJ       abc        R1        1.0
ENDATA
""",
    "error_overflow": """NAME          OVF
ROWS
 N  OBJ
 L  R1
COLUMNS
    X         OBJ       1e999      R1        1.0
ENDATA
""",
    "error_rhs_pairs": """NAME          RHSPAIRS
ROWS
 N  OBJ
 L  R1
 L  R2
COLUMNS
    X         R1        1.0
RHS
    RHS1      R1        8.0        R2
ENDATA
""",
    "error_contradictory_bounds": """NAME          CONTRA
ROWS
 N  OBJ
 L  R1
COLUMNS
    X         R1        1.0
    Y         R1        1.0
BOUNDS
 LO BND       Y         5.0
 UP BND       Y         3.0
ENDATA
""",
}


def parse_full_dump(text: str) -> dict:
    lines = text.split("\n")
    out, i = {}, 0
    while i < len(lines) and lines[i]:
        key, *rest = lines[i].split(" ")
        if key == "name":
            out["name"] = lines[i + 1]
            i += 2
        elif key in ("row_names", "col_names"):
            k = int(rest[0])
            out[key] = lines[i + 1:i + 1 + k]
            i += 1 + k
        elif key in ("sense", "has_quadratic"):
            out[key] = int(rest[0])
            i += 1
        elif key == "obj_offset":
            out[key] = rest[0]
            i += 1
        elif key == "dims":
            out["m"], out["n"], out["nnz"] = (int(v) for v in rest)
            i += 1
        elif key in ("col_ptr", "row_ind", "var_types", "q_col_ptr", "q_row_ind"):
            out[key] = [int(v) for v in rest[1:]]
            i += 1
        else:  # float vectors, kept as the exact hex text
            out[key] = rest[1:]
            i += 1
    return out


def ref(args, td):
    return subprocess.run([str(REFSIMPLEX), *map(str, args)], capture_output=True, text=True,
                          timeout=60, cwd=td)


def main():
    cases = []
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        inputs = [(p.stem, "reference data/toy/" + p.name, p.read_bytes().decode())
                  for p in sorted(TOYS.glob("*.mps"))]
        inputs += [(k, "crafted", v) for k, v in CRAFTED.items()]
        for name, source, text in inputs:
            src = td / f"{name}.mps"
            src.write_bytes(text.encode())
            res = ref(["--mps-dump", src, td / "dump.txt"], td)
            case = dict(name=name, source=source, text=text)
            if res.returncode == 3:
                case["error"] = res.stdout.strip()[len("error "):]
            else:
                res.check_returncode()
                case["parse"] = parse_full_dump((td / "dump.txt").read_text())
                ref(["--mps-write", src, td / "out.mps"], td).check_returncode()
                case["written"] = (td / "out.mps").read_text()
            cases.append(case)
            print(name, "error: " + case["error"] if "error" in case else
                  f"m={case['parse']['m']} n={case['parse']['n']} nnz={case['parse']['nnz']}")
        gens = [("gen_staircase_T1_R20_C48_s3", S.gen_staircase(1, 20, 48, 3)),
                ("gen_random_m12_n30_k3_s7", S.gen_random(12, 30, 3, 7))]
        for name, p in gens:
            ours = td / f"{name}.mps"
            S.write_mps(p, str(ours))
            ref(["--mps-write", ours, td / "out.mps"], td).check_returncode()
            cases.append(dict(name=name, source="sih_solver generator, written by write_mps",
                              text=ours.read_text(), written=(td / "out.mps").read_text()))
            print(name, "written", len(cases[-1]["written"]), "bytes")
    out = Path(__file__).resolve().parent / "mps_fixtures.json"
    out.write_text(json.dumps({"reader": "reference io::read_mps / io::write_mps "
                               "(core/src/io/mps_io.cpp) via oracle/ref/ref_driver.cpp",
                               "cases": cases}, indent=1) + "\n")


if __name__ == "__main__":
    main()
