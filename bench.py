#!/usr/bin/env python3
"""Benchmark of the MI355X PDHG LP engine (one JSON line on stdout).

Workloads (synthetic, seeded):
  large-ref  one random equality LP, m=500,000, n=1,200,000, 10 nnz/column
             (12M nnz), reference-mode PDHG; a step is one PDHG iteration.

The roofline object reports the dominant fused SpMV kernel: algorithmic
bytes per launch divided by its average launch duration from HIP events
recorded inside the timed solve; peak = MI355X HBM3E 8 TB/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "refinery-optimizer_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline_ref(p, iters: int):
    """The oracle (CPU restatement of the reference PDHG, FP32) timed on a
    bounded number of iterations of the same LP, one thread."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    lp = O.LpArrays.of(p)
    t = time.perf_counter()
    O.pdhg_ref(lp, max_iterations=iters, tol_primal=0.0, tol_dual=0.0)
    dt = time.perf_counter() - t
    return iters / dt, dt


def bench_large_ref(args):
    import sih_solver as S

    t = time.perf_counter()
    p = S.gen_random(500_000, 1_200_000, 10, 1)
    log(f"generated m={p.num_rows()} n={p.num_cols()} nnz={p.num_nonzeros()} "
        f"in {time.perf_counter() - t:.1f}s")
    cfg = S.GpuPdhgConfig(max_iterations=max(args.warmup, 1), tol_primal=0.0, tol_dual=0.0)
    S.solve_gpu_pdhg(p, cfg)  # warm-up: code objects, allocator
    cfg = S.GpuPdhgConfig(max_iterations=args.steps, tol_primal=0.0, tol_dual=0.0, profile=True)
    sol = S.solve_gpu_pdhg(p, cfg)
    steps = args.steps
    ms_step = 1e3 * sol.time_solver_sec / steps
    n_l = max(sol.kernel_launches_csc, 1)
    csc_us = 1e3 * sol.kernel_ms_spmv_csc / n_l
    csr_us = 1e3 * sol.kernel_ms_spmv_csr / n_l
    gbs_csc = sol.bytes_csc_launch / (csc_us * 1e-6) / 1e9
    gbs_csr = sol.bytes_csr_launch / (csr_us * 1e-6) / 1e9
    log(f"csc fused {csc_us:.1f} us ({gbs_csc:.0f} GB/s), csr fused {csr_us:.1f} us "
        f"({gbs_csr:.0f} GB/s), {ms_step:.3f} ms/iteration")
    dom = ("csc", csc_us, sol.bytes_csc_launch, gbs_csc) if csc_us >= csr_us else \
        ("csr", csr_us, sol.bytes_csr_launch, gbs_csr)
    out = {
        "metric": "PDHG iterations/sec (reference mode); SpMV HBM GB/s vs peak",
        "value": steps / sol.time_solver_sec,
        "unit": "iterations/s",
        "ms_per_step": ms_step,
        "dtype": "fp32",
        "config": {"workload": "large-ref: random equality LP m=500000 n=1200000 "
                               "10 nnz/col, reference-mode PDHG, step = one iteration"},
        "roofline": {"bound": "hbm", "kernel": f"k_ref_{'primal' if dom[0] == 'csc' else 'dual'}",
                     "achieved": dom[3], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom[3] / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": dom[2], "us_per_launch": dom[1]},
        "spmv_gbs": {"csc_fused": gbs_csc, "csr_fused": gbs_csr},
    }
    if not args.no_cpu:
        it = args.cpu_iters
        v, dt = cpu_baseline_ref(p, it)
        out["cpu_baseline"] = {"value": v, "unit": "iterations/s", "cores": 1, "kind": "port",
                               "sample": f"{it} reference-mode FP32 iterations of the same LP "
                                         f"({dt:.1f}s, oracle, 1 thread)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="large-ref", choices=["large-ref"])
    ap.add_argument("--cpu-iters", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every GPU box of this build has one, and the "
              "engine has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    out = bench_large_ref(args)
    out.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "data": "synthetic"})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
