"""Synthetic workloads the engine is measured on (sizes as named below).

* ``random_equality(m, n, k, seed)``: m equality rows, n columns, k nonzeros
  per column at uniformly random rows, values U[-1, 1], 0 <= x <= 10,
  b = A x0 (x0 ~ U[0, 10]), c = A^T y0 + s (y0 ~ N(0, 1), s ~ U[0, 1] zeroed on
  40% of the columns).  m=1,000 n=2,500 k=8 is the CPU-sized case; m=500,000
  n=1,200,000 k=10 (12M nonzeros) the one-GPU large case.
* ``staircase(T)``: refinery planning LP of T time blocks of 400 rows x 900
  columns (product balances, blend-quality rows, unit-capacity rows; crude,
  sale and inventory columns; inventory links block t to t+1).
* ``scenario_sweep(count)``: `count` LPs with m log-uniform in [200, 20,000]
  and n = 2.4 m, alternating the two generators, each with lognormal(0, 0.3)
  price shocks on every cost; 10% of the scenarios zero the capacity of one
  random unit-capacity row (staircase scenarios carry such rows).
"""
from __future__ import annotations

import math
from typing import List

from . import Problem, gen_random, gen_scenario, gen_staircase


class _SplitMix:
    """SplitMix64: the composition of a sweep (sizes, kinds, seeds)."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def u01(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53


def random_equality(m: int, n: int, k: int, seed: int) -> Problem:
    return gen_random(m, n, k, seed)


def staircase(T: int, seed: int, R: int = 400, C: int = 900) -> Problem:
    return gen_staircase(T, R, C, seed)


def sweep_specs(count: int = 512, m_lo: int = 200, m_hi: int = 20000, seed: int = 2024):
    """(kind, size, base_seed, scenario_seed, zero_capacity) per scenario."""
    g = _SplitMix(seed)
    specs = []
    for s in range(count):
        m = int(math.floor(math.exp(math.log(m_lo) + (math.log(m_hi) - math.log(m_lo)) * g.u01())))
        m = max(m_lo, min(m_hi, m))
        base_seed, scn_seed = g.next(), g.next()
        zero_cap = g.u01() < 0.1
        if s % 2 == 0:
            specs.append(("random", (m, int(round(2.4 * m)), 8), base_seed, scn_seed, zero_cap))
        else:
            T = max(1, min(96, int(round(m / 400))))
            R = max(10, int(round(m / T)))
            specs.append(("staircase", (T, R, int(round(2.4 * R))), base_seed, scn_seed, zero_cap))
    return specs


def scenario_sweep(count: int = 512, m_lo: int = 200, m_hi: int = 20000, seed: int = 2024,
                   shock_sigma: float = 0.3) -> List[Problem]:
    return [gen_scenario(kind, size, bs, ss, shock_sigma, zc)
            for kind, size, bs, ss, zc in sweep_specs(count, m_lo, m_hi, seed)]
