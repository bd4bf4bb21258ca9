#!/usr/bin/env python3
"""Benchmark of the MI355X PDHG LP engine (one JSON line on stdout).

Default workload (``--workload sweep512``): 512 refinery scenario LPs, m
log-uniform in [200, 20,000], n = 2.4 m (alternating random-equality and
staircase generators, lognormal(0, 0.3) price shocks, 10% with one unit
capacity zeroed), ~40.6M nonzeros in total.  A step is one batched solve of
all 512 LPs to FP64 relative KKT <= 1e-6 (FP32 iterates, FP64 checks and
iterative refinement).  value = LPs that reached 1e-6 / device-resident
solve time of the batch.  The same engine solving the LPs one per launch
sequence ("plain") is timed on every 8th LP of the sweep and reported beside
it, as is the CPU port of the same algorithm on a bounded sample.

Other workloads:
  large-ref  one random equality LP m=500,000 n=1,200,000 10 nnz/column,
             reference-mode PDHG; a step is one PDHG iteration.
  large-enh  the same LP, enhanced mode to 1e-4 relative KKT.
  copy       a 256 MiB device-to-device elementwise copy (torch), the
             calibration kernel for the FETCH_SIZE / WRITE_SIZE counters.

The roofline object reports the dominant fused SpMV kernel: algorithmic
bytes per launch / its average launch duration from HIP events recorded on
the solve's stream inside the timed solve; peak = MI355X HBM3E 8 TB/s.
`traffic` is the HBM bytes per launch of that kernel measured with the
FETCH_SIZE and WRITE_SIZE counters (two rocprofv3 --pmc passes over
`bench.py --max-iterations 128`, every instance active), scaled by what the
counters read for the copy workload; the figures are read from
profiles/pmc_traffic_sweep512.md, which records the commands.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "refinery-optimizer_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "LPs/sec to 1e-6 rel KKT, batched vs plain GPU PDHG; SpMV HBM GB/s vs peak"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic_sweep512.md"


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed PMC summary, or None.

    The summary holds one line `traffic_bytes <kernel> <bytes>` per kernel
    (FETCH_SIZE x fetch factor + WRITE_SIZE x write factor, in bytes)."""
    if not TRAFFIC_FILE.exists():
        return None
    m = re.search(rf"^traffic_bytes {re.escape(kernel)} (\d+)\s*$", TRAFFIC_FILE.read_text(),
                  re.MULTILINE)
    return float(m.group(1)) if m else None


def roofline(sol, kernels=("k_enh_primal", "k_enh_dual"), traffic=True):
    n_l = max(sol.kernel_launches_csc, 1)
    csc_us = 1e3 * sol.kernel_ms_spmv_csc / n_l
    csr_us = 1e3 * sol.kernel_ms_spmv_csr / n_l
    gbs_csc = sol.bytes_csc_launch / (csc_us * 1e-6) / 1e9 if csc_us > 0 else 0.0
    gbs_csr = sol.bytes_csr_launch / (csr_us * 1e-6) / 1e9 if csr_us > 0 else 0.0
    if csc_us >= csr_us:
        name, us, b, g = kernels[0], csc_us, sol.bytes_csc_launch, gbs_csc
    else:
        name, us, b, g = kernels[1], csr_us, sol.bytes_csr_launch, gbs_csr
    return ({"bound": "hbm", "kernel": name, "achieved": g, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": g / HBM_PEAK_GBS, "traffic": measured_traffic(name) if traffic else None,
             "bytes_per_launch": b, "us_per_launch": us},
            {"csc_fused_gbs": gbs_csc, "csr_fused_gbs": gbs_csr, "csc_fused_us": csc_us,
             "csr_fused_us": csr_us})


def cpu_port_sample(probs, tol, threads):
    """The CPU port of the same algorithm (oracle) on a sample, `threads` LPs at a time."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    def one(p):
        r = O.pdhg_enh(O.LpArrays.of(p), tol=tol, max_iterations=200000, max_refinements=8,
                       refine_ratio=1e-2)
        return r["kkt"] <= tol

    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        ok = list(ex.map(one, probs))
    return sum(ok), time.perf_counter() - t


def bench_sweep(args):
    import sih_solver as S
    from sih_solver import workloads as W

    tol = 1e-6
    t = time.perf_counter()
    probs = W.scenario_sweep(512)
    nnz = sum(p.num_nonzeros() for p in probs)
    log(f"generated 512 LPs ({nnz} nonzeros) in {time.perf_counter() - t:.1f}s")
    cfg = S.GpuPdhgConfig.enhanced(tol=tol, max_iterations=args.max_iterations,
                                   launch_graph=0 if args.no_graph else 1)
    for w in range(args.warmup):
        S.solve_batch(probs, cfg)
        log(f"warmup step {w} done")
    times, walls, conv = [], [], 0
    sols = None
    for s in range(args.steps):
        cfg.profile = (s == args.steps - 1)
        sols = S.solve_batch(probs, cfg)
        times.append(sols[0].time_solver_sec)
        walls.append(sols[0].time_wall_sec)
        conv = sum(1 for x in sols if x.status == S.SolutionStatus.Optimal and x.rel_kkt <= tol)
        log(f"step {s}: {conv}/512 at <= {tol:g} in {times[-1]:.3f}s device loop "
            f"({walls[-1]:.3f}s with host setup), max iters {max(x.simplex_iterations for x in sols)}")
    dev = sum(times) / len(times)
    wall = sum(walls) / len(walls)
    rf, spmv = roofline(sols[0])
    # plain: the same engine, one LP per launch sequence, on every 8th LP
    sample = probs[::8] if not args.no_plain else []
    cfg.profile = False
    t_plain, ok_plain = 0.0, 0
    for p in sample:
        s1 = S.solve_gpu_pdhg(p, cfg)
        t_plain += s1.time_solver_sec
        ok_plain += int(s1.status == S.SolutionStatus.Optimal and s1.rel_kkt <= tol)
    plain_rate = ok_plain / t_plain if t_plain > 0 and ok_plain else None
    if plain_rate:
        log(f"plain: {ok_plain}/{len(sample)} in {t_plain:.3f}s -> {plain_rate:.1f} LPs/s; "
            f"batched {conv / dev:.1f} LPs/s ({conv / dev / plain_rate:.2f}x)")
    capped = args.max_iterations < 100000
    out = {
        "metric": METRIC, "value": conv / dev, "unit": "LPs/s", "ms_per_step": 1e3 * dev,
        "dtype": "fp32 iterates, fp64 KKT/refinement",
        "config": {"workload": "sweep512: 512 scenario LPs, m log-uniform [200, 20000], n=2.4m, "
                               f"{nnz} nnz, FP64 rel. KKT <= 1e-6, step = one batched solve"
                               + (f" (profiling run capped at {args.max_iterations} iterations: "
                                  "not a result)" if capped else ""),
                   "lps": 512, "tol": tol, "max_iterations": args.max_iterations},
        "converged": conv,
        "wall_lps_per_s": conv / wall,
        "plain": None if plain_rate is None else {
            "lps_per_s": plain_rate, "sample": f"every 8th LP ({len(sample)} LPs), one solve "
                                               "call each", "converged": ok_plain,
            "seconds": t_plain},
        "batched_over_plain": (conv / dev) / plain_rate if plain_rate else None,
        "roofline": rf, "spmv": spmv,
    }
    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        # every 16th LP of the sweep with m <= 4000 (bounded CPU time: the
        # largest LPs alone would take minutes on the host cores)
        cs = [p for p in probs[::16] if p.num_rows() <= 4000][: args.cpu_lps]
        ok, dt = cpu_port_sample(cs, tol, cores)
        out["cpu_baseline"] = {"value": ok / dt, "unit": "LPs/s", "cores": cores, "kind": "port",
                               "sample": f"{len(cs)} LPs of the sweep (every 16th with m <= 4000), CPU port of the "
                                         f"same algorithm, {cores} LPs in parallel, {dt:.1f}s, "
                                         f"{ok} converged"}
    return out


def bench_large(args, mode):
    import sih_solver as S

    t = time.perf_counter()
    p = S.gen_random(500_000, 1_200_000, 10, 1)
    log(f"generated m={p.num_rows()} n={p.num_cols()} nnz={p.num_nonzeros()} "
        f"in {time.perf_counter() - t:.1f}s")
    if mode == "ref":
        cfg = S.GpuPdhgConfig(max_iterations=max(args.warmup, 1), tol_primal=0.0, tol_dual=0.0)
        S.solve_gpu_pdhg(p, cfg)
        cfg = S.GpuPdhgConfig(max_iterations=args.steps, tol_primal=0.0, tol_dual=0.0, profile=True)
        sol = S.solve_gpu_pdhg(p, cfg)
        rf, spmv = roofline(sol, ("k_ref_primal", "k_ref_dual"))
        out = {"metric": "PDHG iterations/s (reference mode); SpMV HBM GB/s vs peak",
               "value": args.steps / sol.time_solver_sec, "unit": "iterations/s",
               "ms_per_step": 1e3 * sol.time_solver_sec / args.steps, "dtype": "fp32",
               "config": {"workload": "large-ref: random equality LP m=500000 n=1200000 10 nnz/col, "
                                      "reference-mode PDHG, step = one iteration"}}
        # the same iterations with rocSPARSE products (bench/cusparse_pdhg_baseline.cu counterpart)
        from sih_solver import vendor

        out["vendor_rocsparse"] = {}
        for name, mode_v in (("transpose_op", vendor.TRANSPOSE_OP), ("explicit_csc", vendor.EXPLICIT_CSC)):
            v = vendor.run(p, args.steps, mode_v)
            out["vendor_rocsparse"][name] = {"us_per_iter": v.us_per_iter, "us_spmv_t": v.us_spmv_t,
                                             "us_spmv": v.us_spmv}
            log(f"rocSPARSE {name}: {v.us_per_iter:.1f} us/iter (A^T y {v.us_spmv_t:.1f} us, "
                f"A x {v.us_spmv:.1f} us); engine {1e6 * sol.time_solver_sec / args.steps:.1f} us/iter")
    else:
        cfg = S.GpuPdhgConfig.enhanced(tol=1e-4, max_iterations=args.steps, profile=True)
        sol = S.solve_gpu_pdhg(p, cfg)
        rf, spmv = roofline(sol)
        out = {"metric": "time to 1e-4 rel KKT (enhanced); SpMV HBM GB/s vs peak",
               "value": sol.time_solver_sec, "unit": "s", "higher_is_better": False,
               "ms_per_step": 1e3 * sol.time_solver_sec / max(sol.simplex_iterations, 1),
               "dtype": "fp32 iterates, fp64 KKT", "status": int(sol.status),
               "rel_kkt": sol.rel_kkt, "iterations": sol.simplex_iterations,
               "config": {"workload": "large-enh: random equality LP m=500000 n=1200000 10 nnz/col "
                                      "to 1e-4 rel KKT, step = one iteration"}}
    out["roofline"] = rf
    out["spmv"] = spmv
    return out


def bench_copy(args):
    """256 MiB float32 device copy: reads 256 MiB and writes 256 MiB per launch."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the copy workload needs a HIP device")
    n = 64 << 20
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    gbs = 2 * 4 * n * args.steps / dt / 1e9
    return {"metric": "device copy GB/s (counter calibration)", "value": gbs, "unit": "GB/s",
            "ms_per_step": 1e3 * dt / args.steps, "dtype": "f32",
            "config": {"workload": "copy: 256 MiB read + 256 MiB write per launch"},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": 2.0 * 4 * n}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sweep512",
                    choices=["sweep512", "large-ref", "large-enh", "copy"])
    ap.add_argument("--max-iterations", type=int, default=100000,
                    help="sweep512 iteration cap (a small cap makes a short profiling run in "
                         "which every instance stays active; its value is then not a result)")
    ap.add_argument("--cpu-lps", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-LP-per-call comparison")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernel by kernel instead of replaying recorded check cycles "
                         "(the kernels are the same; rocprofv3's kernel tracing faults on graph "
                         "replays in this ROCm release, so profiling runs use this)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every GPU box of this build has one, and the "
              "engine has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "sweep512":
        out = bench_sweep(args)
    elif args.workload == "copy":
        out = bench_copy(args)
    else:
        out = bench_large(args, "ref" if args.workload == "large-ref" else "enh")
    out.setdefault("higher_is_better", True)
    out.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "scaling": "weak",
                "vs_baseline": None, "data": "synthetic"})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
