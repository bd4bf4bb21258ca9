"""Scenario tooling on the GPU PDHG path.

* ``standard_perturbations()``: five operational perturbations of a refinery
  staircase LP, after py/refinery/reoptimization.py:33-78 (crude price shock,
  product demand surge, unit turndown, sweet-crude supply cut, a market move
  in prices and demands), expressed as scenario options of
  ``gen_scenario_ex`` (the perturbations of
  py/refinery/parameterized_generator.py:28-133 on the staircase model).
* ``PdhgReoptimizer``: RefineryReoptimizer.run_benchmark
  (reoptimization.py:80-158) with the GPU PDHG in place of the dual simplex:
  the base LP is solved, then each perturbation is solved cold and warm, and
  cold/warm iteration counts, times and objectives are recorded under the
  reference's keys.  Each perturbation is applied to the base LP, as there;
  the warm start is the base solution's x, y and step size.  The reference
  instead cascades its simplex basis from one perturbation's solution to the
  next (:154-156); for PDHG that start is farther from the next scenario's
  solution (on a 4-block staircase: 35,200 warm iterations in total against
  44,288 cold, while the base solution needs 22,592), so ``cascade=True``
  keeps it as an option.  Reusing the base's primal weight as well
  (``reuse_primal_weight=True``) cost iterations on two of the five
  perturbations there (25,344 in total; cascaded with it, 53,760).
"""
from __future__ import annotations

import time
from typing import Any, Dict, List, Optional, Sequence

from . import GpuPdhgConfig, Problem, SolutionStatus, WarmStart, gen_scenario_ex, gen_staircase, solve_batch


def standard_perturbations() -> List[Dict[str, Any]]:
    """Five perturbations (reoptimization.py:33-78) as gen_scenario_ex options."""
    return [
        {"id": "pert_1_crude_price", "name": "Crude Price Shock",
         "desc": "crude prices +30% (high-crude regime) with a 5% spread",
         "opts": {"regime": "high_crude", "price_std": 0.05}},
        {"id": "pert_2_product_demand", "name": "Product Demand Surge",
         "desc": "product prices +30%, diesel products +25% more with their demand caps raised",
         "opts": {"regime": "diesel_surge"}},
        {"id": "pert_3_unit_turndown", "name": "Unit Turndown",
         "desc": "first process unit of each period at 65% capacity (turnaround regime)",
         "opts": {"regime": "turnaround"}},
        {"id": "pert_4_sweet_shortage", "name": "Sweet Crude Supply Cut",
         "desc": "second crude's supply at 35% (sweet-crude shortage regime)",
         "opts": {"regime": "sweet_crude_shortage"}},
        {"id": "pert_5_market_move", "name": "Market Move",
         "desc": "prices, demands, supplies and capacities perturbed (std 0.10 / 0.15 / 0.15 / 0.08)",
         "opts": {"price_std": 0.10, "demand_std": 0.15, "supply_std": 0.15, "capacity_std": 0.08}},
    ]


class PdhgReoptimizer:
    """Cold vs warm GPU PDHG re-solves of perturbed refinery staircase LPs."""

    def __init__(self, T: int = 12, R: int = 400, C: int = 900, base_seed: int = 1,
                 config: Optional[GpuPdhgConfig] = None):
        self.size = (T, R, C)
        self.base_seed = base_seed
        self.config = config or GpuPdhgConfig.enhanced(tol=1e-6)

    def base_problem(self) -> Problem:
        return gen_staircase(*self.size, seed=self.base_seed)

    def scenario(self, index: int, opts: Dict[str, Any]) -> Problem:
        return gen_scenario_ex("staircase", self.size, self.base_seed, 1000 + index, **opts)

    def run_benchmark(self, perturbations: Optional[Sequence[Dict[str, Any]]] = None, cascade: bool = False,
                      reuse_primal_weight: bool = False) -> List[Dict[str, Any]]:
        """Solve the base LP, then every perturbation cold and warm.  The warm
        start is the base solution (x, y, step size; with reuse_primal_weight
        its primal weight too), or with cascade=True the previous
        perturbation's warm solution, as reoptimization.py:104-156 does with
        simplex bases."""
        perturbations = list(perturbations or standard_perturbations())
        cfg = self.config
        base = solve_batch([self.base_problem()], cfg)[0]
        if base.status != SolutionStatus.Optimal:
            raise RuntimeError(f"base LP not solved (status {base.status.name})")

        def start(sol):
            return WarmStart(sol.x, sol.row_duals, sol.primal_weight if reuse_primal_weight else 0.0,
                             sol.step_size)

        current = start(base)
        results = []
        for k, pert in enumerate(perturbations):
            p = self.scenario(k, pert["opts"])
            t0 = time.perf_counter()
            cold = solve_batch([p], cfg)[0]
            t_cold = 1e3 * (time.perf_counter() - t0)
            t0 = time.perf_counter()
            warm = solve_batch([p], cfg, [current])[0]
            t_warm = 1e3 * (time.perf_counter() - t0)
            ci, wi = cold.simplex_iterations, warm.simplex_iterations
            results.append({
                "id": pert["id"], "name": pert["name"], "desc": pert["desc"],
                "cold_iters": ci, "warm_iters": wi,
                "iter_reduction_pct": 100.0 * (ci - wi) / ci if ci > 0 else 0.0,
                "cold_time_ms": t_cold, "warm_time_ms": t_warm,
                "cold_device_ms": 1e3 * cold.time_solver_sec, "warm_device_ms": 1e3 * warm.time_solver_sec,
                "speedup": t_cold / t_warm if t_warm > 1e-4 else 1.0,
                "objective": warm.primal_objective, "cold_objective": cold.primal_objective,
                "cold_status": cold.status.name, "warm_status": warm.status.name,
                "warm_rel_kkt": warm.rel_kkt, "cold_rel_kkt": cold.rel_kkt,
            })
            if cascade and warm.status == SolutionStatus.Optimal:
                current = start(warm)
        return results

    @staticmethod
    def format_results_table(results: List[Dict[str, Any]]) -> str:
        """Markdown table with the reference's columns (reoptimization.py:160-180)."""
        lines = [
            "| Scenario / Perturbation | Cold Iters | Warm Iters | Iter Reduction (%) | Cold Time (ms) "
            "| Warm Time (ms) | Speedup | Objective |",
            "| :--- | :--- | :--- | :--- | :--- | :--- | :--- | :--- |",
        ]
        for r in results:
            lines.append(
                f"| **{r['name']}** ({r['desc']}) | {r['cold_iters']} | **{r['warm_iters']}** | "
                f"**{r['iter_reduction_pct']:.1f}%** | {r['cold_time_ms']:.2f} ms | "
                f"**{r['warm_time_ms']:.2f} ms** | **{r['speedup']:.2f}x** | {r['objective']:,.2f} |")
        return "\n".join(lines)
