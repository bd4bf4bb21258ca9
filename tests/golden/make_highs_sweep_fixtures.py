"""Objective values of a two-family sample of the 512-LP scenario sweep from
scipy's HiGHS (FP64).

The sweep (sih_solver.workloads.scenario_sweep(512)) alternates the random
equality family (even indices) and the refinery staircase family (odd
indices).  The sample takes both families across their size range: HiGHS
solves staircase LPs quickly up to m ~ 6,000 rows, while random-family LPs
get slow above m ~ 1,000 (m = 1,107 takes ~20 s, m ~ 2,000 several
minutes), so the random-family part stops there.  The LPs are regenerated
from their sweep specs by the engine's host generator (no GPU needed).

    python tests/golden/make_highs_sweep_fixtures.py   # writes highs_sweep.json
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
from sih_solver import workloads as W  # noqa: E402
from test_oracle import _highs  # noqa: E402

# sweep indices: random family (m 203 .. 1,107), staircase family (m 212 .. 6,090)
RANDOM = [56, 232, 108, 504, 110, 478, 498]
STAIRCASE = [207, 347, 147, 315, 459, 369, 97, 331, 275, 507, 9, 483]


def main():
    out = HERE / "highs_sweep.json"
    done = json.loads(out.read_text()) if out.exists() else []
    have = {c["index"] for c in done}
    specs = W.sweep_specs(512)
    for q in RANDOM + STAIRCASE:
        if q in have:
            continue
        kind, size, bs, ss, zc = specs[q]
        p = S.gen_scenario(kind, size, bs, ss, 0.3, zc)
        t = time.time()
        s, _ = _highs(O.LpArrays.of(p))
        done.append(dict(index=q, kind=kind, m=p.num_rows(), n=p.num_cols(), nnz=p.num_nonzeros(),
                         zero_capacity=bool(zc), objective=float(s.fun), solver="scipy HiGHS",
                         seconds=round(time.time() - t, 2)))
        done.sort(key=lambda c: c["index"])
        out.write_text(json.dumps(done, indent=1))
        print(done[-1] if done[-1]["index"] == q else next(c for c in done if c["index"] == q), flush=True)


if __name__ == "__main__":
    main()
