"""Objective values of full-size staircase LPs from scipy's HiGHS (FP64).

The reference FP64 dual simplex cannot solve these sizes in reasonable time
(its LU refactorisation is O(m^3), sparse_lu.cpp:238-251), so the error of
the engine at full size is stated against HiGHS; at small T the reference
simplex fixtures (make_fixtures.py) pin it.  The LPs are regenerated from
their seeds by the engine's host generator (no GPU needed).

    python tests/golden/make_highs_fixtures.py   # writes highs_staircase.json
"""
import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
for p in (ROOT / "refinery-optimizer_amd", ROOT / "oracle", ROOT / "tests"):
    sys.path.insert(0, str(p))

import oracle as O  # noqa: E402
import sih_solver as S  # noqa: E402
from test_oracle import _highs  # noqa: E402

CASES = [(96, 400, 900, 1), (48, 400, 900, 1), (12, 400, 900, 1)]


def main():
    out = HERE / "highs_staircase.json"
    done = json.loads(out.read_text()) if out.exists() else []
    have = {(c["T"], c["R"], c["C"], c["seed"]) for c in done}
    for T, R, Cb, seed in CASES:
        if (T, R, Cb, seed) in have:
            continue
        p = S.gen_staircase(T, R, Cb, seed)
        t = time.time()
        s, _ = _highs(O.LpArrays.of(p))
        done.append(dict(T=T, R=R, C=Cb, seed=seed, m=p.num_rows(), n=p.num_cols(),
                         nnz=p.num_nonzeros(), objective=float(s.fun), solver="scipy HiGHS",
                         seconds=round(time.time() - t, 1)))
        out.write_text(json.dumps(done, indent=1))
        print(done[-1], flush=True)


if __name__ == "__main__":
    main()
