#!/usr/bin/env python3
"""Benchmark of the MI355X PDHG LP engine (one JSON line on stdout).

Default workload (``--workload sweep512``): 512 refinery scenario LPs, m
log-uniform in [200, 20,000], n = 2.4 m (alternating random-equality and
staircase generators, lognormal(0, 0.3) price shocks, 10% with one unit
capacity zeroed), ~40.6M nonzeros in total.  A step is one batched solve of
all 512 LPs to FP64 relative KKT <= 1e-6 (FP32 iterates, FP64 checks and
iterative refinement).  value = LPs that reached 1e-6 / device-resident
solve time of the batch.  The same engine solving the LPs one per launch
sequence ("plain") is timed on every 16th LP of each generator family and
reported beside it, as is the CPU port of the same algorithm on a bounded
two-family sample (one core and all host cores).

Other workloads:
  cfg0       the CPU-sized random equality LP m=1,000 n=2,500 8 nnz/column:
             reference mode to the reference's own stop (P, D <= 1e-4) on
             the GPU beside the same algorithm in FP64 on one host core.
  large-ref  one random equality LP m=500,000 n=1,200,000 10 nnz/column,
             reference-mode PDHG; a step is one PDHG iteration.
  large-enh  the same LP, enhanced mode to 1e-4 relative KKT.
  staircase96  one refinery staircase LP (96 blocks of 400 x 900) to 1e-8,
             objective error against the committed FP64 HiGHS value.
  adjoint64  64 staircase LPs (48 blocks) forward to 1e-8 and backward
             through the torch.autograd layer, with central finite
             differences on 4 of them at their FP64 crossover vertices.
  hybrid48   solve_hybrid on a 48-block staircase LP: GPU PDHG to 1e-8,
             then the FP64 crossover on the host, against HiGHS's optimum.
  copy       a 256 MiB device-to-device elementwise copy (torch), the
             calibration kernel for the FETCH_SIZE / WRITE_SIZE counters.
  stream     the device's streaming ceilings: float4 kernels reading 1 GiB
             (cached and non-temporal) and copying it, best of five grids.
  gather     the fused kernels' memory pattern without their arithmetic: the
             index/value stream of 40.6M nonzeros in 8,192-nonzero chunks,
             alone and with sorted gathers 0.25 ... 6 vector entries apart.
  longrow    enhanced-mode iterations on a random LP with one 500k-nonzero
             row and one 100k-nonzero column (the row-split rule).

The roofline object reports the dominant fused SpMV kernel: algorithmic
bytes per launch (the fused-iteration formula 16 nnz + 32 n + 24 m, split by
launch; see fused_bytes) / its average launch duration from HIP events
recorded on the solve's stream inside the timed solve; peak = MI355X HBM3E
8 TB/s.  `engine_achieved` uses the engine's own per-launch count, which adds
the enhanced mode's Halpern anchor reads.
`traffic` is the HBM bytes per launch of that kernel measured with the
FETCH_SIZE and WRITE_SIZE counters (two rocprofv3 --pmc passes over
`bench.py --max-iterations 128`, every instance active), scaled by what the
counters read for the copy workload; the figures are read from
profiles/pmc_traffic_sweep512.md, which records the commands.

`vs_fp64_vertex` (sweep512) states the error of the batched results against
FP64 optimal vertices: after the timed solves, the crossover (mpdhg_polish,
host code) starts from each sampled LP's PDHG point and reaches a vertex;
the relative objective difference is reported (max and median, overall and
per family).
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


def measured_traffic(kernel: str, summary: Path = TRAFFIC_FILE):
    """HBM bytes per launch of `kernel` from a committed PMC summary of its
    workload (the sweep's by default), or None.

    The summary holds one line `traffic_bytes <kernel> <bytes>` per kernel
    (FETCH_SIZE x fetch factor + WRITE_SIZE x write factor, in bytes)."""
    if summary is None or not summary.exists():
        return None
    m = re.search(rf"^traffic_bytes {re.escape(kernel)} (\d+)\s*$", summary.read_text(), re.MULTILINE)
    return float(m.group(1)) if m else None


# The fused enhanced kernels read the matrix stream non-temporally (the
# k_enh_*_nt instantiations) while the listed instances' stream, both
# directions, exceeds the 256 MiB last-level cache.
NT_STREAM_BYTES = 256 * 2**20


def enh_kernels(nnz: int):
    sfx = "_nt" if 16 * nnz > NT_STREAM_BYTES else ""
    return ("k_enh_primal" + sfx, "k_enh_dual" + sfx)


def fused_bytes(probs, vector_steps):
    """Algorithmic bytes of one fused primal launch (A^T y and the primal
    step) and one fused dual launch (A xbar and the dual step) over the LPs
    `probs`, by the fused-iteration formula 16 nnz + 32 n + 24 m (scalar
    steps; +4 n + 4 m with per-column / per-row steps) split by launch:
    primal 8 nnz + 4 (n + 1) col pointers + 4 m gathered y + 6 n vector words
    (c, l, u, x read; x, xbar written; + tau with vector steps); dual 8 nnz +
    4 (m + 1) row pointers + 4 n gathered xbar + 4 m vector words (rl, ru, y
    read; y written; + sigma).  The enhanced mode's Halpern anchor reads (x0,
    y0) and its extrapolated y are real traffic outside this formula; the
    engine's own per-launch figure counts them."""
    nnz = sum(p.num_nonzeros() for p in probs)
    n = sum(p.num_cols() for p in probs)
    m = sum(p.num_rows() for p in probs)
    v = 4 if vector_steps else 0
    return (8.0 * nnz + 4.0 * (n + len(probs)) + 4.0 * m + (24.0 + v) * n,
            8.0 * nnz + 4.0 * (m + len(probs)) + 4.0 * n + (16.0 + v) * m)


def roofline(sol, kernels=("k_enh_primal", "k_enh_dual"), traffic=TRAFFIC_FILE, algo=None):
    """The dominant fused kernel's rate.  algo = (primal, dual) algorithmic
    bytes per launch (fused_bytes) gives `achieved`; the engine's own count
    per launch (sol.bytes_*_launch) is reported beside it."""
    n_l = max(sol.kernel_launches_csc, 1)
    csc_us = 1e3 * sol.kernel_ms_spmv_csc / n_l
    csr_us = 1e3 * sol.kernel_ms_spmv_csr / n_l
    a_csc, a_csr = algo if algo else (sol.bytes_csc_launch, sol.bytes_csr_launch)
    gbs_csc = a_csc / (csc_us * 1e-6) / 1e9 if csc_us > 0 else 0.0
    gbs_csr = a_csr / (csr_us * 1e-6) / 1e9 if csr_us > 0 else 0.0
    if csc_us >= csr_us:
        name, us, b, e = kernels[0], csc_us, a_csc, sol.bytes_csc_launch
    else:
        name, us, b, e = kernels[1], csr_us, a_csr, sol.bytes_csr_launch
    g = b / (us * 1e-6) / 1e9 if us > 0 else 0.0
    ge = e / (us * 1e-6) / 1e9 if us > 0 else 0.0
    return ({"bound": "hbm", "kernel": name, "achieved": g, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": g / HBM_PEAK_GBS, "traffic": measured_traffic(name, traffic) if traffic else None,
             "bytes_per_launch": b, "us_per_launch": us,
             "engine_bytes_per_launch": e, "engine_achieved": ge, "engine_frac": ge / HBM_PEAK_GBS},
            {"csc_fused_gbs": gbs_csc, "csr_fused_gbs": gbs_csr, "csc_fused_us": csc_us,
             "csr_fused_us": csr_us,
             # tile schedules (threads per workgroup, nonzeros per LDS chunk) the solve used,
             # and the time the schedule search took
             "csr_schedule": [sol.csr_block, sol.csr_chunk], "csc_schedule": [sol.csc_block, sol.csc_chunk],
             "schedule_search_ms": sol.tune_ms})


def cpu_port_sample(probs, tol, threads):
    """The CPU port of the same algorithm (oracle) on a sample, `threads` LPs at a time."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    def one(p):
        r = O.pdhg_enh(O.LpArrays.of(p), tol=tol, max_iterations=200000, max_refinements=8,
                       refine_ratio=1e-2)
        return r["kkt"] <= tol

    t = time.perf_counter()
    if threads == 1:
        ok = [one(p) for p in probs]
    else:
        with ThreadPoolExecutor(max_workers=threads) as ex:
            ok = list(ex.map(one, probs))
    return sum(ok), time.perf_counter() - t


def host_cores():
    """Cores this process may use (at most 16, the host share of one GPU) and the host's CPU model."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return min(16, avail), {"nproc": os.cpu_count(), "affinity": avail, "model": model}


# The sweep alternates its two generator families (random equality LPs at
# even indices, staircase LPs at odd ones), so every 16th LP of each family
# samples both families over their whole size ranges.
def strata(count=512, step=16):
    return sorted(list(range(0, count, step)) + list(range(1, count, step)))


# Size caps of the vertex check: the crossover's LU fills in on random-family
# bases (m = 1,914 takes ~6 s on the host), while staircase bases stay sparse.
VERTEX_M_MAX = {"random": 1200, "staircase": 8000}


def vertex_check(probs, sols, tol):
    """Objective error of the batched PDHG results against FP64 optimal vertices:
    the crossover (mpdhg_polish from each PDHG point, host code, outside the
    timed region) on every 8th LP of each family up to the family's size cap."""
    import numpy as np

    import sih_solver as S

    from sih_solver import workloads as W

    family = [sp[0] for sp in W.sweep_specs(len(probs))]
    picks = [q for q in strata(len(probs), 8) if probs[q].num_rows() <= VERTEX_M_MAX[family[q]]]
    errs = {"random": [], "staircase": []}
    t0, opt = time.perf_counter(), 0
    for q in picks:
        p, s = probs[q], sols[q]
        cb, rb = S.identify_basis(p, s.x)
        r = S.polish(p, cb, rb, s.x)
        if r.status != S.SolutionStatus.Optimal:
            continue
        opt += 1
        errs[family[q]].append(abs(s.primal_objective - r.objective) / max(1.0, abs(r.objective)))
    dt = time.perf_counter() - t0
    allv = errs["random"] + errs["staircase"]
    log(f"vertex check: {opt}/{len(picks)} FP64 vertices in {dt:.1f}s, objective error "
        f"max {max(allv):.2e}, median {float(np.median(allv)):.2e}")
    fam = {k: {"lps": len(v), "max_rel_objective_error": max(v) if v else None,
               "median_rel_objective_error": float(np.median(v)) if v else None}
           for k, v in errs.items()}
    return {"sample": f"every 8th LP of each family, random family m <= {VERTEX_M_MAX['random']}, "
                      f"staircase family m <= {VERTEX_M_MAX['staircase']} ({len(picks)} LPs), FP64 "
                      "crossover from the batched PDHG point", "vertices": opt,
            "max_rel_objective_error": max(allv), "median_rel_objective_error": float(np.median(allv)),
            "per_family": fam, "pdhg_tol": tol, "seconds": dt}


def bench_sweep(args):
    import sih_solver as S
    from sih_solver import workloads as W

    tol = 1e-6
    t = time.perf_counter()
    probs = W.scenario_sweep(512)
    nnz = sum(p.num_nonzeros() for p in probs)
    log(f"generated 512 LPs ({nnz} nonzeros) in {time.perf_counter() - t:.1f}s")
    cfg = S.GpuPdhgConfig.enhanced(tol=tol, max_iterations=args.max_iterations,
                                   launch_graph=0 if args.no_graph else 1, spmv_tune=args.spmv_tune,
                                   spmv_block=args.spmv_block, spmv_chunk=args.spmv_chunk)
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
    rf, spmv = roofline(sols[0], enh_kernels(nnz), algo=fused_bytes(probs, False))
    # plain: the same engine, one LP per launch sequence, on every 16th LP of
    # each family (64 LPs, half of each family)
    family = [sp[0] for sp in W.sweep_specs(len(probs))]
    sample = strata(len(probs), 16) if not args.no_plain else []
    cfg.profile = False
    fam_t = {"random": 0.0, "staircase": 0.0}
    fam_ok = {"random": 0, "staircase": 0}
    for q in sample:
        s1 = S.solve_gpu_pdhg(probs[q], cfg)
        fam_t[family[q]] += s1.time_solver_sec
        fam_ok[family[q]] += int(s1.status == S.SolutionStatus.Optimal and s1.rel_kkt <= tol)
    t_plain, ok_plain = sum(fam_t.values()), sum(fam_ok.values())
    plain_rate = ok_plain / t_plain if t_plain > 0 and ok_plain else None
    if plain_rate:
        log(f"plain: {ok_plain}/{len(sample)} in {t_plain:.3f}s -> {plain_rate:.1f} LPs/s "
            f"(random {fam_ok['random'] / max(fam_t['random'], 1e-9):.1f}, staircase "
            f"{fam_ok['staircase'] / max(fam_t['staircase'], 1e-9):.1f}); "
            f"batched {conv / dev:.1f} LPs/s ({conv / dev / plain_rate:.2f}x)")
    vertex = None if args.no_vertex else vertex_check(probs, sols, tol)
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
            "lps_per_s": plain_rate, "sample": f"every 16th LP of each family ({len(sample)} LPs, "
                                               "half random-equality, half staircase), one solve "
                                               "call each", "converged": ok_plain,
            "seconds": t_plain,
            "per_family": {k: {"lps_per_s": fam_ok[k] / fam_t[k] if fam_t[k] > 0 else None,
                               "converged": fam_ok[k], "seconds": fam_t[k]} for k in fam_t}},
        "batched_over_plain": (conv / dev) / plain_rate if plain_rate else None,
        "roofline": rf, "spmv": spmv,
        "vs_fp64_vertex": vertex,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_sweep(probs, family, cfg, tol, args)
    return out


def cpu_baseline_sweep(probs, family, cfg, tol, args):
    """The CPU port of the same algorithm on a bounded two-family sample of the
    sweep: every 16th LP of each family with m <= 2,000 (24 LPs, ~30 s on one
    core), timed on one core and on all the host cores (LPs in parallel), with
    the batched GPU solve of the same LPs beside it.  (The largest sweep LPs
    alone would take minutes each on the host.)"""
    import sih_solver as S

    idx = [q for q in strata(len(probs), 16) if probs[q].num_rows() <= 2000][: args.cpu_lps]
    cs = [probs[q] for q in idx]
    nf = {k: sum(1 for q in idx if family[q] == k) for k in ("random", "staircase")}
    cores, host = host_cores()
    ok1, dt1 = cpu_port_sample(cs, tol, 1)
    okn, dtn = cpu_port_sample(cs, tol, cores)
    g = S.solve_batch(cs, cfg)
    okg = sum(1 for x in g if x.status == S.SolutionStatus.Optimal and x.rel_kkt <= tol)
    log(f"cpu port: {len(cs)} LPs ({nf['random']} random, {nf['staircase']} staircase) "
        f"{ok1 / dt1:.2f} LPs/s on 1 core ({dt1:.1f}s), {okn / dtn:.2f} LPs/s on {cores} cores "
        f"({dtn:.1f}s); GPU batched {okg / g[0].time_solver_sec:.1f} LPs/s on the same LPs")
    return {"value": okn / dtn, "unit": "LPs/s", "cores": cores, "kind": "port",
            "sample": f"{len(cs)} LPs of the sweep: every 16th LP of each family with m <= 2000 "
                      f"({nf['random']} random-equality, {nf['staircase']} staircase), CPU port of the "
                      f"same algorithm (FP32 iterates, FP64 checks) to 1e-6, {cores} LPs in parallel: "
                      f"{dtn:.1f}s, {okn} converged",
            "single_core": {"value": ok1 / dt1, "unit": "LPs/s", "seconds": dt1, "converged": ok1},
            "gpu_same_sample": {"value": okg / g[0].time_solver_sec, "unit": "LPs/s",
                                "seconds": g[0].time_solver_sec, "converged": okg,
                                "note": "batched solve of the same LPs, device-resident loop"},
            "host": host}


def bench_large(args, mode):
    import sih_solver as S

    t = time.perf_counter()
    p = S.gen_random(500_000, 1_200_000, 10, 1)
    log(f"generated m={p.num_rows()} n={p.num_cols()} nnz={p.num_nonzeros()} "
        f"in {time.perf_counter() - t:.1f}s")
    if mode == "ref":
        cfg = S.GpuPdhgConfig(max_iterations=max(args.warmup, 1), tol_primal=0.0, tol_dual=0.0,
                              spmv_bands=args.spmv_bands)
        S.solve_gpu_pdhg(p, cfg)
        cfg = S.GpuPdhgConfig(max_iterations=args.steps, tol_primal=0.0, tol_dual=0.0,
                              spmv_tune=args.spmv_tune, launch_graph=0 if args.no_graph else 1,
                              spmv_bands=args.spmv_bands)
        sol = S.solve_gpu_pdhg(p, cfg)
        # kernel timings from a separate run whose first cycles launch eagerly with events
        cfg.profile, cfg.max_iterations = True, min(args.steps, 100)
        kd = "k_ref_dual" if sol.csr_bands == 1 else f"k_ref_dual_band x{sol.csr_bands}"
        rf, spmv = roofline(S.solve_gpu_pdhg(p, cfg), ("k_ref_primal", kd),
                            traffic=ROOT / "profiles" / "pmc_traffic_large_ref.md",
                            algo=fused_bytes([p], True))
        out = {"metric": "PDHG iterations/s (reference mode); SpMV HBM GB/s vs peak",
               "value": args.steps / sol.time_solver_sec, "unit": "iterations/s",
               "ms_per_step": 1e3 * sol.time_solver_sec / args.steps, "dtype": "fp32",
               "config": {"workload": "large-ref: random equality LP m=500000 n=1200000 10 nnz/col, "
                                      "reference-mode PDHG, step = one iteration"}}
        # the same iterations with rocSPARSE products (bench/cusparse_pdhg_baseline.cu counterpart)
        from sih_solver import vendor

        out["csr_bands"] = sol.csr_bands
        out["vendor_rocsparse"] = {}
        for name, mode_v in () if args.no_vendor else (("transpose_op", vendor.TRANSPOSE_OP),
                                                       ("explicit_csc", vendor.EXPLICIT_CSC)):
            v = vendor.run(p, args.steps, mode_v)
            out["vendor_rocsparse"][name] = {"us_per_iter": v.us_per_iter, "us_spmv_t": v.us_spmv_t,
                                             "us_spmv": v.us_spmv}
            log(f"rocSPARSE {name}: {v.us_per_iter:.1f} us/iter (A^T y {v.us_spmv_t:.1f} us, "
                f"A x {v.us_spmv:.1f} us); engine {1e6 * sol.time_solver_sec / args.steps:.1f} us/iter")
    else:
        cfg = S.GpuPdhgConfig.enhanced(tol=1e-4, max_iterations=args.steps, profile=True,
                                       spmv_tune=args.spmv_tune, launch_graph=0 if args.no_graph else 1,
                                       spmv_bands=args.spmv_bands)
        sol = S.solve_gpu_pdhg(p, cfg)
        kp, kd = enh_kernels(p.num_nonzeros())
        if sol.csr_bands > 1:  # the dual product's time spans its band launches
            kd = kd.replace("k_enh_dual", "k_enh_dual_band") + f" x{sol.csr_bands}"
        rf, spmv = roofline(sol, (kp, kd), traffic=ROOT / "profiles" / "pmc_traffic_large_enh.md",
                            algo=fused_bytes([p], False))
        out = {"metric": "time to 1e-4 rel KKT (enhanced); SpMV HBM GB/s vs peak",
               "value": sol.time_solver_sec, "unit": "s", "higher_is_better": False,
               "ms_per_step": 1e3 * sol.time_solver_sec / max(sol.simplex_iterations, 1),
               "dtype": "fp32 iterates, fp64 KKT", "status": int(sol.status),
               "rel_kkt": sol.rel_kkt, "iterations": sol.simplex_iterations,
               "config": {"workload": "large-enh: random equality LP m=500000 n=1200000 10 nnz/col "
                                      "to 1e-4 rel KKT, step = one iteration"},
               "csr_bands": sol.csr_bands}
    out["roofline"] = rf
    out["spmv"] = spmv
    return out


def bench_cfg0(args):
    """The CPU-sized random equality LP (m=1,000, n=2,500, 8 nonzeros per
    column, seed 0).  The reference's algorithm to its own stop (P, D <= 1e-4,
    gpu_solver.cpp:252-270) on the GPU in reference mode (FP32), beside the
    FP64 CPU path: the same algorithm in double on one host core (the
    oracle's restatement; the reference has no CPU PDHG of its own).  The
    enhanced mode to 1e-4 relative KKT is reported too."""
    import numpy as np

    import sih_solver as S

    p = S.gen_random(1000, 2500, 8, 0)
    K = 100000
    cfg = S.GpuPdhgConfig(max_iterations=K)
    for _ in range(args.warmup):
        S.solve_gpu_pdhg(p, cfg)
    ts = []
    for _ in range(max(args.steps, 1)):
        sol = S.solve_gpu_pdhg(p, cfg)
        ts.append(sol.time_solver_sec)
    dev = float(np.median(ts))
    log(f"cfg0 GPU reference mode: status {int(sol.status)} in {sol.simplex_iterations} iterations, "
        f"{dev * 1e3:.1f} ms (P {sol.ref_primal:.2e}, D {sol.ref_dual:.2e}, rel KKT {sol.rel_kkt:.2e})")
    ecfg = S.GpuPdhgConfig.enhanced(tol=1e-4)
    S.solve_gpu_pdhg(p, ecfg)
    enh = S.solve_gpu_pdhg(p, ecfg)
    log(f"cfg0 GPU enhanced mode: rel KKT {enh.rel_kkt:.2e} in {enh.simplex_iterations} iterations, "
        f"{enh.time_solver_sec * 1e3:.1f} ms")
    out = {"metric": "seconds to the reference's stop (P, D <= 1e-4), config-0 LP",
           "value": dev, "unit": "s", "higher_is_better": False, "ms_per_step": 1e3 * dev,
           "dtype": "fp32",
           "config": {"workload": "cfg0: random equality LP m=1000 n=2500 8 nnz/col seed 0, "
                                  "reference-mode PDHG to P, D <= 1e-4; step = one solve"},
           "status": int(sol.status), "iterations": sol.simplex_iterations,
           "ref_primal": sol.ref_primal, "ref_dual": sol.ref_dual, "rel_kkt": sol.rel_kkt,
           "enhanced": {"seconds": enh.time_solver_sec, "iterations": enh.simplex_iterations,
                        "rel_kkt": enh.rel_kkt, "status": int(enh.status), "tol": 1e-4}}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O

        lp = O.LpArrays.of(p)
        O.pdhg_ref(lp, max_iterations=K, fp64=True)  # warm-up
        runs = []
        for _ in range(5):
            t = time.perf_counter()
            r = O.pdhg_ref(lp, max_iterations=K, fp64=True)
            runs.append(time.perf_counter() - t)
        k = O.kkt(lp, r["x"], r["y"])
        cpu = float(np.median(runs))
        log(f"cfg0 CPU FP64 path: status {r['status']} in {r['iterations']} iterations, "
            f"{cpu * 1e3:.1f} ms median of 5 (rel KKT {k['kkt']:.2e})")
        _, host = host_cores()
        out["cpu_baseline"] = {"value": cpu, "unit": "s", "cores": 1, "kind": "port",
                               "sample": "the whole config-0 solve: the reference algorithm in FP64 on "
                                         "one host core (int32 CSC/CSR, sequential row sums) to P, D <= "
                                         f"1e-4; median of 5 after a warm-up; status {r['status']}, "
                                         f"{r['iterations']} iterations, rel KKT {k['kkt']:.2e}",
                               "host": host}
    return out


HIGHS_FIXTURES = ROOT / "tests" / "golden" / "highs_staircase.json"


def highs_objective(T, R, Cb, seed):
    """FP64 objective of a staircase LP from the committed HiGHS fixtures, or None."""
    if not HIGHS_FIXTURES.exists():
        return None
    for c in json.loads(HIGHS_FIXTURES.read_text()):
        if (c["T"], c["R"], c["C"], c["seed"]) == (T, R, Cb, seed):
            return c["objective"]
    return None


def bench_staircase(args):
    """Refinery staircase LP, 96 time blocks of 400 rows x 900 columns, to 1e-8."""
    import sih_solver as S

    T, R, Cb, seed, tol = 96, 400, 900, 1, 1e-8
    p = S.gen_staircase(T, R, Cb, seed)
    log(f"staircase T={T}: m={p.num_rows()} n={p.num_cols()} nnz={p.num_nonzeros()}")
    cfg = S.GpuPdhgConfig.enhanced(tol=tol, max_iterations=400000,
                                   launch_graph=0 if args.no_graph else 1, spmv_tune=args.spmv_tune)
    for _ in range(args.warmup):
        S.solve_gpu_pdhg(p, cfg)
    times = []
    for s in range(args.steps):
        cfg.profile = (s == args.steps - 1)
        sol = S.solve_gpu_pdhg(p, cfg)
        times.append(sol.time_solver_sec)
        log(f"step {s}: status {int(sol.status)} kkt {sol.rel_kkt:.2e} in {sol.simplex_iterations} "
            f"iterations, {times[-1]:.3f}s")
    ref = highs_objective(T, R, Cb, seed)
    rf, spmv = roofline(sol, enh_kernels(p.num_nonzeros()), traffic=False, algo=fused_bytes([p], False))
    dev = sum(times) / len(times)
    return {"metric": "seconds to 1e-8 rel KKT (staircase LP, FP32 iterates + FP64 refinement)",
            "value": dev, "unit": "s", "higher_is_better": False, "ms_per_step": 1e3 * dev,
            "dtype": "fp32 iterates, fp64 KKT/refinement",
            "config": {"workload": f"staircase{T}: T={T} blocks of {R}x{Cb}, m={p.num_rows()} "
                                   f"n={p.num_cols()} nnz={p.num_nonzeros()}, seed {seed}, to {tol:g}"},
            "status": int(sol.status), "iterations": sol.simplex_iterations,
            "restarts": sol.restarts, "refinements": sol.refinements, "rel_kkt": sol.rel_kkt,
            "objective": sol.primal_objective, "highs_objective": ref,
            "rel_objective_error_vs_highs": (abs(sol.primal_objective - ref) / max(1.0, abs(ref))
                                             if ref is not None else None),
            "roofline": rf, "spmv": spmv}


def bench_hybrid(args):
    """GpuSolver::solve_hybrid on a 48-block staircase LP (19,200 x 43,200; --blocks
    12 or 96 for the other sizes with HiGHS fixtures):
    enhanced PDHG on the GPU to 1e-8, the basis hand-off, and the FP64
    crossover (primal push from the PDHG point, then the simplex) on the
    host.  Reported beside the HiGHS optimum and HiGHS's own solve time,
    both from the committed fixture (the reference's simplex factorises in
    O(m^3) and cannot run at this size)."""
    import sih_solver as S

    T, R, Cb, seed, tol = args.blocks, 400, 900, 1, 1e-8
    p = S.gen_staircase(T, R, Cb, seed)
    log(f"hybrid staircase T={T}: m={p.num_rows()} n={p.num_cols()} nnz={p.num_nonzeros()}")
    cfg = S.GpuPdhgConfig.enhanced(tol=tol, max_iterations=400000, spmv_tune=args.spmv_tune)
    for _ in range(args.warmup):
        S.solve_gpu_hybrid(p, cfg)
    walls = []
    for s in range(args.steps):
        sol = S.solve_gpu_hybrid(p, cfg)
        walls.append(sol.time_wall_sec)
        log(f"step {s}: status {int(sol.status)} objective {sol.primal_objective:.12g}; PDHG "
            f"{sol.time_solver_sec:.3f}s on the device (kkt {sol.rel_kkt:.1e}), whole call "
            f"{sol.time_wall_sec:.3f}s, {sol.simplex_iterations} simplex iterations")
    ref = next((c for c in json.loads(HIGHS_FIXTURES.read_text())
                if (c["T"], c["R"], c["C"], c["seed"]) == (T, R, Cb, seed)), None)
    wall = sum(walls) / len(walls)
    return {"metric": "seconds to an optimal FP64 vertex (GPU PDHG + host crossover)",
            "value": wall, "unit": "s", "higher_is_better": False, "ms_per_step": 1e3 * wall,
            "dtype": "fp32 PDHG iterates, fp64 KKT and simplex",
            "config": {"workload": f"hybrid staircase{T}: m={p.num_rows()} n={p.num_cols()} "
                                   f"nnz={p.num_nonzeros()}, seed {seed}, PDHG to {tol:g}"},
            "status": int(sol.status), "objective": sol.primal_objective,
            "pdhg_seconds": sol.time_solver_sec, "pdhg_rel_kkt": sol.rel_kkt,
            "crossover_seconds": sol.time_wall_sec - sol.time_solver_sec,
            "simplex_iterations": sol.simplex_iterations,
            "highs_objective": ref["objective"] if ref else None,
            "highs_seconds": ref["seconds"] if ref else None,
            "rel_objective_error_vs_highs": (abs(sol.primal_objective - ref["objective"]) /
                                             abs(ref["objective"]) if ref else None)}


def bench_adjoint(args):
    """64 staircase LPs (48 blocks each): batched forward to 1e-8 through the
    torch.autograd layer, the adjoint backward, and central finite differences
    of L = w . x*(b) along random unit directions on 4 of the LPs."""
    import numpy as np
    import torch

    import sih_solver as S
    from sih_solver.autograd import _shifted, lp_layer

    K, T, R, Cb, tol = 64, 48, 400, 900, 1e-8
    t = time.perf_counter()
    probs = [S.gen_staircase(T, R, Cb, 100 + k) for k in range(K)]
    nnz = sum(p.num_nonzeros() for p in probs)
    log(f"generated {K} staircase LPs T={T} ({nnz} nonzeros) in {time.perf_counter() - t:.1f}s")
    cfg = S.GpuPdhgConfig.enhanced(tol=tol, max_iterations=1000000, launch_graph=0 if args.no_graph else 1)
    rng = np.random.default_rng(7)
    c = torch.tensor(np.concatenate([p.c for p in probs]), dtype=torch.float64, requires_grad=True)
    b = torch.zeros(sum(p.num_rows() for p in probs), dtype=torch.float64, requires_grad=True)
    w = torch.tensor(rng.normal(size=c.numel()))
    t0 = time.perf_counter()
    x = lp_layer(probs, c, b, cfg)
    t1 = time.perf_counter()
    (w * x).sum().backward()
    t2 = time.perf_counter()
    log(f"forward {t1 - t0:.2f}s, backward {t2 - t1:.2f}s")
    # Central differences at FP64 vertices of the 48-block LPs.  A PDHG point
    # lies inside the optimal face, where x*(b) is set-valued; at a vertex
    # with its optimal basis, x*(b) is the basis's solution B^-1 (b - N x_N)
    # as long as that basis stays optimal.  solve_hybrid (PDHG to 1e-9, the
    # basis hand-off, the FP64 crossover) gives the vertex and basis of each
    # sampled LP; the GPU adjoint is evaluated there, and L = w . x* is
    # differenced along d = A_RB v (B: basic columns inside their bounds, R:
    # rows at a bound), a shift that keeps the basis primal feasible and so
    # optimal: each shifted LP is re-solved by the FP64 simplex from that
    # basis (no pivots expected).  These vertices are dual degenerate (a few
    # hundred nonbasic columns with zero reduced cost), so other optimal
    # vertices exist: re-solving from scratch could land on another one.
    import scipy.sparse as sp

    hcfg = S.GpuPdhgConfig.enhanced(tol=1e-9, max_iterations=2000000)
    co = np.concatenate([[0], np.cumsum([p.num_cols() for p in probs])])
    wn = w.numpy()
    errs, steps, screened, checks, t_fd = [], [], [], [], time.perf_counter()
    for k in range(min(8, K)):
        if len(errs) == args.adjoint_fd:
            break
        p = probs[k]
        wk = wn[co[k]:co[k + 1]]
        v0 = S.solve_gpu_hybrid(p, hcfg)
        if v0.status != S.SolutionStatus.Optimal:
            screened.append(dict(lp=k, reason=f"crossover status {v0.status.name}"))
            continue
        xk, yk = v0.x, -v0.row_duals  # simplex sign -> PDHG sign
        r = S.adjoint_batch([p], [xk], [yk], [wk])[0]
        A = sp.csc_matrix((p.A.csc_values(), p.A.csc_row_ind(), p.A.csc_col_ptr()),
                          shape=(p.num_rows(), p.num_cols()))
        ax = A @ xk
        isb = (xk - p.col_lower > 1e-7 * (1 + abs(xk))) & (p.col_upper - xk > 1e-7 * (1 + abs(xk)))
        fin = np.concatenate([p.row_lower[np.isfinite(p.row_lower)], p.row_upper[np.isfinite(p.row_upper)]])
        tl = 1e-7 * (1 + np.abs(fin).max(initial=0.0))
        isr = ((np.abs(yk) > 1e-7 * (1 + np.abs(yk).max())) | (np.abs(ax - p.row_lower) <= tl) |
               (np.abs(ax - p.row_upper) <= tl))
        B, R = np.flatnonzero(isb), np.flatnonzero(isr)
        AB = A[R][:, B]
        sres = np.linalg.norm(AB.T @ r.dl_db[R] - wk[B]) / np.linalg.norm(wk[B])
        dual_deg = int(np.sum(np.abs(v0.reduced_costs[~isb]) <= 1e-9 * (1 + np.abs(p.c).max())))
        if sres > 1e-6:
            screened.append(dict(lp=k, system_residual=sres, zero_reduced_costs=dual_deg))
            log(f"LP {k}: vertex |B|={B.size} |R|={R.size}: ||A_RB^T mu - w_B||/||w_B|| = {sres:.1e}: "
                "A_RB short of column rank, no finite difference")
            continue
        cb = np.array([int(z) for z in v0.col_basis], np.int8)
        rb = np.array([int(z) for z in v0.row_basis], np.int8)
        v = rng.normal(size=B.size)
        d = np.zeros(p.num_rows())
        d[R] = AB @ v
        scale = np.linalg.norm(d)
        d /= scale
        v /= scale  # a unit shift h d moves x_B by h v
        room = np.minimum(xk[B] - p.col_lower[B], p.col_upper[B] - xk[B])
        hs = [1e-4 * max(1.0, np.abs(fin).max(initial=0.0)),
              0.5 * float(np.min(room / np.maximum(np.abs(v), 1e-300)))]
        dv = A[~isr][:, B] @ v
        lo_gap, hi_gap = (ax - p.row_lower)[~isr], (p.row_upper - ax)[~isr]
        with np.errstate(divide="ignore", invalid="ignore"):
            lim = np.where(dv < 0, lo_gap / -dv, np.where(dv > 0, hi_gap / dv, np.inf))
        hs.append(0.5 * float(np.min(lim, initial=np.inf)))
        h = min(hs)
        xs, pivots = [], []
        for sgn in (1.0, -1.0):
            vs = S.polish(_shifted(p, p.c, sgn * h * d), cb, rb)
            xs.append(vs.x if vs.status == S.SolutionStatus.Optimal else None)
            pivots.append(vs.iterations)
        if xs[0] is None or xs[1] is None:
            screened.append(dict(lp=k, reason="a shifted LP did not end optimal from the basis"))
            continue
        fd = (wk @ xs[0] - wk @ xs[1]) / (2 * h)
        an = r.dl_db @ d
        errs.append(abs(fd - an) / max(abs(fd), abs(an), 1e-12))
        steps.append(h)
        checks.append(dict(lp=k, basic_cols=int(B.size), active_rows=int(R.size), h=h, fd=fd, adjoint=an,
                           rel_error=errs[-1], system_residual=sres, zero_reduced_costs=dual_deg,
                           shifted_pivots=pivots))
        log(f"LP {k}: vertex |B|={B.size} |R|={R.size} ({dual_deg} zero reduced costs) h={h:.3g} FD {fd:.9e} "
            f"adjoint {an:.9e} rel err {errs[-1]:.2e}; ||A_RB^T mu - w_B||/||w_B|| = {sres:.2e}; "
            f"pivots from the basis {pivots}")
    t_fd = time.perf_counter() - t_fd
    vert = adjoint_fd_at_vertices(rng)
    return {"metric": "batched forward + adjoint seconds (64 staircase LPs to 1e-8)",
            "value": t2 - t0, "unit": "s", "higher_is_better": False, "ms_per_step": 1e3 * (t2 - t0),
            "dtype": "fp32 iterates, fp64 KKT/refinement and adjoint",
            "config": {"workload": f"adjoint64: {K} staircase LPs T={T} ({R}x{Cb} blocks), {nnz} nnz, "
                                   "L = w.x*, w ~ N(0,1), forward to 1e-8"},
            "forward_s": t1 - t0, "backward_s": t2 - t1,
            "dl_dc_all_zero": bool(torch.all(c.grad == 0)),
            "fd_max_rel_error": max(errs) if errs else None, "fd_checks": len(errs),
            "fd_at": "FP64 crossover vertices of the 48-block LPs (solve_hybrid, PDHG to 1e-9); shifted "
                     "LPs re-solved by the FP64 simplex from the vertex's basis",
            "fd_target": 1e-4, "fd_steps": steps, "fd_seconds": t_fd, "fd_detail": checks,
            "not_unique_vertex": screened, "fd_at_vertices_4_blocks": vert}


def _highs_solve(p, rl, ru):
    """FP64 HiGHS solution (x, y in the PDHG sign y = -dobj/db) of p with row bounds rl, ru."""
    import numpy as np
    import scipy.sparse as sp
    from scipy.optimize import linprog

    A = sp.csc_matrix((p.A.csc_values(), p.A.csc_row_ind(), p.A.csc_col_ptr()),
                      shape=(p.num_rows(), p.num_cols()))
    eq = rl == ru
    fu, fl = (~eq) & np.isfinite(ru), (~eq) & np.isfinite(rl)
    Aub = sp.vstack([A[fu], -A[fl]]).tocsr()
    bub = np.concatenate([ru[fu], -rl[fl]])
    s = linprog(p.c, A_ub=Aub if Aub.shape[0] else None, b_ub=bub if bub.size else None,
                A_eq=A[eq] if eq.any() else None, b_eq=rl[eq] if eq.any() else None,
                bounds=list(zip(p.col_lower, np.where(np.isfinite(p.col_upper), p.col_upper, None))),
                method="highs")
    if s.status != 0:
        return None, None, A
    y = np.zeros(p.num_rows())
    if eq.any():
        y[eq] = -s.eqlin.marginals
    mu = -s.ineqlin.marginals if bub.size else np.zeros(0)
    y[fu] += mu[: int(fu.sum())]
    y[fl] -= mu[int(fu.sum()):]
    return s.x, y, A


def adjoint_fd_at_vertices(rng, seeds=(100, 101, 102, 103, 104, 105), T=4):
    """Central differences where x*(b) is a differentiable function: at the
    optimal vertices (FP64 HiGHS) of 4-block staircase LPs.  The GPU adjoint
    is evaluated at the vertex; the direction d = A_RB v keeps the active set
    and the step is 1e-4 of the largest finite row bound.  LPs whose vertex is
    primal degenerate in a way that makes A_RB rank deficient are skipped."""
    import numpy as np

    import sih_solver as S

    errs, used = [], []
    for seed in seeds:
        p = S.gen_staircase(T, 400, 900, seed)
        x, y, A = _highs_solve(p, p.row_lower, p.row_upper)
        if x is None:
            continue
        w = rng.normal(size=p.num_cols())
        r = S.adjoint_batch([p], [x], [y], [w])[0]
        ax = A @ x
        isb = (x - p.col_lower > 1e-7 * (1 + abs(x))) & (p.col_upper - x > 1e-7 * (1 + abs(x)))
        fin = np.concatenate([p.row_lower[np.isfinite(p.row_lower)], p.row_upper[np.isfinite(p.row_upper)]])
        tl = 1e-7 * (1 + np.abs(fin).max())
        isr = (np.abs(y) > 1e-7 * (1 + np.abs(y).max())) | (np.abs(ax - p.row_lower) <= tl) | \
              (np.abs(ax - p.row_upper) <= tl)
        B, R = np.flatnonzero(isb), np.flatnonzero(isr)
        AB = A[R][:, B]
        if np.linalg.norm(AB.T @ r.dl_db[R] - w[B]) > 1e-6 * np.linalg.norm(w[B]):
            continue  # A_RB rank deficient at this vertex: the derivative is not unique
        v = rng.normal(size=B.size)
        d = np.zeros(p.num_rows())
        d[R] = AB @ v
        d /= np.linalg.norm(d)
        h = 1e-4 * max(1.0, np.abs(fin).max())
        xp, _, _ = _highs_solve(p, p.row_lower + h * d, p.row_upper + h * d)
        xm, _, _ = _highs_solve(p, p.row_lower - h * d, p.row_upper - h * d)
        if xp is None or xm is None:
            continue
        fd = (w @ xp - w @ xm) / (2 * h)
        an = r.dl_db @ d
        errs.append(abs(fd - an) / max(abs(fd), abs(an), 1e-12))
        used.append(seed)
        log(f"vertex FD: staircase T={T} seed {seed}: |B|={B.size} |R|={R.size} FD {fd:.9e} "
            f"adjoint {an:.9e} rel err {errs[-1]:.1e}")
    return {"lps": f"staircase T={T} (400x900 blocks), seeds {used}, HiGHS FP64 vertices",
            "checks": len(errs), "max_rel_error": max(errs) if errs else None}


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


def longrow_lp(seed: int = 3):
    """A random equality LP (m = 100,000, n = 500,000, 4 nonzeros per column at
    random rows, values U[-1, 1]) plus one dense row (row 0 over every column)
    and one dense column (column 0 over every row): 2,599,950 nonzeros.
    b = A x0 (x0 ~ U[0, 10]), 0 <= x <= 10, c = A^T y0 + s."""
    import numpy as np
    import scipy.sparse as sp

    import sih_solver as S

    rng = np.random.default_rng(seed)
    m, n = 100000, 500000
    rows = np.sort(rng.integers(0, m, size=(n, 4)), axis=1).ravel()
    A = sp.csc_matrix((rng.uniform(-1, 1, 4 * n), rows, np.arange(0, 4 * n + 1, 4)), shape=(m, n)).tolil()
    A[0, :] = rng.uniform(-1, 1, n)
    A[:, 0] = rng.uniform(-1, 1, (m, 1))
    A = A.tocsc()
    A.sum_duplicates()
    A.sort_indices()
    p = S.Problem("longrow")
    p.resize(m, n)
    p.A = S.SparseMatrix.from_csc(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data)
    x0 = rng.uniform(0, 10, n)
    b = A @ x0
    p.c = A.T @ rng.normal(size=m) + rng.uniform(0, 1, n)
    p.row_lower, p.row_upper = b.copy(), b.copy()
    p.col_lower, p.col_upper = np.zeros(n), np.full(n, 10.0)
    return p


def bench_longrow(args):
    """Enhanced-mode iterations on an LP with a dense row and a dense column
    (the row-split rule: block sums over many workgroups, in-order sums)."""
    import sih_solver as S

    t = time.perf_counter()
    p = longrow_lp()
    log(f"generated m={p.num_rows()} n={p.num_cols()} nnz={p.num_nonzeros()} in {time.perf_counter() - t:.1f}s")
    K = max(args.steps, 64)
    cfg = S.GpuPdhgConfig.enhanced(tol=0.0, max_iterations=K, max_refinements=0,
                                   launch_graph=0 if args.no_graph else 1)
    for _ in range(args.warmup):
        S.solve_gpu_pdhg(p, cfg)
    sol = S.solve_gpu_pdhg(p, cfg)
    us = 1e6 * sol.time_solver_sec / K
    return {"metric": "enhanced PDHG iterations/s, LP with a 500k-nonzero row and a 100k-nonzero column",
            "value": K / sol.time_solver_sec, "unit": "iterations/s", "ms_per_step": us / 1e3,
            "dtype": "fp32 iterates, fp64 KKT",
            "config": {"workload": f"longrow: random equality LP m=100000 n=500000 4 nnz/col plus a dense row "
                                   f"and a dense column ({p.num_nonzeros()} nnz), enhanced mode, {K} iterations; "
                                   "step = one iteration"},
            "us_per_iteration": us, "rel_kkt": sol.rel_kkt}


def bench_stream(args):
    """Streaming ceilings of the device (bench-only vendor library): best GB/s
    of float4 kernels reading 1 GiB through the caches, reading it
    non-temporally, and copying it -- the practical HBM bandwidth against
    which the fused kernels' roofline fractions are read."""
    from sih_solver import vendor

    r = vendor.stream(1 << 30, max(args.steps, 1))
    return {"metric": "device streaming GB/s (HBM ceilings)", "value": r["read_nt_gbs"], "unit": "GB/s",
            "ms_per_step": 1e3 * (1 << 30) / (r["read_nt_gbs"] * 1e9), "dtype": "f32",
            "config": {"workload": "stream: 1 GiB float4 read (cached, non-temporal) and copy, "
                                   "grids of 1024-16384 workgroups of 256, best grid"},
            "read_gbs": r["read_gbs"], "read_nt_gbs": r["read_nt_gbs"], "copy_gbs": r["copy_gbs"],
            "roofline": {"bound": "hbm", "achieved": r["read_nt_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": r["read_nt_gbs"] / HBM_PEAK_GBS, "traffic": float(1 << 30)}}


# Chunk strides (vector entries between neighbouring sorted gathers) probed
# by the gather workload: the staircase family's chunks sit near 0.25, the
# random family's from 1 (small LPs) to 6 (m = 20,000).
GATHER_STRIDES = (0.25, 1.0, 2.0, 4.0, 6.0)


def bench_gather(args):
    """Gather cost in the fused kernels' memory pattern (bench-only vendor
    library): the 40.6M-nonzero stream of 4,960 chunks of 8,192 alone, then
    with sorted gathers at each stride, then with the LDS pairs and row walk;
    what the stream and the gathers each cost the dual kernel."""
    from sih_solver import vendor

    rows = []
    for st in GATHER_STRIDES:
        r = vendor.gather(st, reps=max(args.steps, 1) * 10)
        rows.append({"stride": st, **r})
        log(f"stride {st:5.2f}: stream {r['stream_us']:.1f} us ({r['stream_gbs']:.0f} GB/s), + gathers "
            f"{r['gather_us']:.1f} us, + LDS pairs and walk {r['staged_us']:.1f} us")
    s0 = rows[0]
    return {"metric": "index/value stream GB/s of the fused kernels' pattern; gather cost by stride",
            "value": s0["stream_gbs"], "unit": "GB/s", "ms_per_step": 1e-3 * s0["stream_us"],
            "dtype": "f32",
            "config": {"workload": "gather: 4,960 chunks of 8,192 (int32 index, FP32 value) words, one "
                                   "512-thread workgroup per chunk, XCD-contiguous chunks, sorted gathers "
                                   f"at strides {list(GATHER_STRIDES)}"},
            "strides": rows,
            "roofline": {"bound": "hbm", "achieved": s0["stream_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": s0["stream_gbs"] / HBM_PEAK_GBS, "traffic": None}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sweep512",
                    choices=["sweep512", "cfg0", "large-ref", "large-enh", "staircase96", "adjoint64",
                             "hybrid48", "copy", "stream", "gather", "longrow"])
    ap.add_argument("--max-iterations", type=int, default=100000,
                    help="sweep512 iteration cap (a small cap makes a short profiling run in "
                         "which every instance stays active; its value is then not a result)")
    ap.add_argument("--cpu-lps", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-LP-per-call comparison")
    ap.add_argument("--no-vertex", action="store_true",
                    help="skip the FP64-vertex objective check of the sweep's results")
    ap.add_argument("--spmv-tune", type=int, default=1, choices=(0, 1, 2),
                    help="tile-schedule search (mpdhg_config.spmv_tune): 1 = time the schedules on "
                         "large batches, 0 = the rule (deterministic choice, used by profiling runs)")
    ap.add_argument("--spmv-block", type=int, default=0, choices=(0, 128, 256, 512),
                    help="sweep512: force the tile workgroup size of both directions (0 = auto)")
    ap.add_argument("--spmv-chunk", type=int, default=0, choices=(0, 1024, 2048, 4096, 8192),
                    help="sweep512: force the tile chunk of both directions (0 = auto)")
    ap.add_argument("--spmv-bands", type=int, default=0,
                    help="large-ref / large-enh: column bands of the A xbar product (mpdhg_config.spmv_bands: "
                         "0 = rule, 1 = none, k = k bands)")
    ap.add_argument("--no-vendor", action="store_true", help="large-ref: skip the rocSPARSE baseline")
    ap.add_argument("--adjoint-fd", type=int, default=4,
                    help="adjoint64: LPs differenced at their FP64 vertices")
    ap.add_argument("--blocks", type=int, default=48, choices=(12, 48, 96),
                    help="hybrid48: time blocks of the staircase LP (HiGHS fixtures exist for these)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernel by kernel instead of replaying recorded check cycles "
                         "(the kernels are the same; rocprofv3's kernel tracing faults on graph "
                         "replays in this ROCm release, so profiling runs use this)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (how the batch and a "
              "single large LP would shard is a design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "sweep512":
        out = bench_sweep(args)
    elif args.workload == "copy":
        out = bench_copy(args)
    elif args.workload == "stream":
        out = bench_stream(args)
    elif args.workload == "gather":
        out = bench_gather(args)
    elif args.workload == "longrow":
        out = bench_longrow(args)
    elif args.workload == "cfg0":
        out = bench_cfg0(args)
    elif args.workload == "staircase96":
        out = bench_staircase(args)
    elif args.workload == "adjoint64":
        out = bench_adjoint(args)
    elif args.workload == "hybrid48":
        out = bench_hybrid(args)
    else:
        out = bench_large(args, "ref" if args.workload == "large-ref" else "enh")
    out.setdefault("higher_is_better", True)
    out.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "scaling": "weak",
                "vs_baseline": None, "data": "synthetic"})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
