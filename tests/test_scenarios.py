"""Scenario generator (mpdhg_gen_scenario_ex, host code) on the staircase model.

The perturbations restate py/refinery/parameterized_generator.py:28-133:
each multiplier lies in the reference's clip range, the regime factors are
exact, and nothing else of the LP changes.  A staircase block of R rows and C
columns holds P = R/2 balance rows, Q = 3R/10 blend rows and U = R - P - Q
capacity rows, and NC = C - 2P crude, P sale and P inventory columns.
"""
import numpy as np
import pytest

import sih_solver as S

T, R, Cb = 3, 40, 90
P, Q = R // 2, (3 * R) // 10
U, NC = R - P - Q, Cb - 2 * P


def kinds():
    col = np.zeros(T * Cb, np.int8)
    row = np.zeros(T * R, np.int8)
    for t in range(T):
        col[t * Cb + NC:t * Cb + NC + P] = 1
        col[t * Cb + NC + P:(t + 1) * Cb] = 2
        row[t * R + P:t * R + P + Q] = 1
        row[t * R + P + Q:(t + 1) * R] = 2
    return col, row


def scen(**kw):
    return S.gen_scenario_ex("staircase", (T, R, Cb), 4, 77, **kw)


BASE = S.gen_staircase(T, R, Cb, 4)
COL, ROW = kinds()


def test_structure_matches_the_generator():
    assert np.array_equal(BASE.row_kind, ROW)
    assert np.all(BASE.c[COL == 0] > 0) and np.all(BASE.c[COL == 1] < 0) and np.all(BASE.c[COL == 2] > 0)


def test_no_perturbation_is_the_base_lp():
    p = scen()
    for k in ("c", "col_lower", "col_upper", "row_lower", "row_upper"):
        assert np.array_equal(getattr(p, k), getattr(BASE, k)), k
    assert np.array_equal(p.A.csc_values(), BASE.A.csc_values())


@pytest.mark.parametrize("regime,target,factor", [
    ("high_crude", "crude", 1.30), ("turnaround", "cap0", 0.65), ("sweet_crude_shortage", "supply1", 0.35)])
def test_regime_factors_are_exact(regime, target, factor):
    p = scen(regime=regime)
    if target == "crude":
        assert np.array_equal(p.c[COL == 0], BASE.c[COL == 0] * factor)
        assert np.array_equal(p.c[COL != 0], BASE.c[COL != 0])
    elif target == "cap0":
        first = np.array([t * R + P + Q for t in range(T)])
        assert np.array_equal(p.row_upper[first], BASE.row_upper[first] * factor)
        rest = np.setdiff1d(np.arange(T * R), first)
        assert np.array_equal(p.row_upper[rest], BASE.row_upper[rest])
    else:
        second = np.array([t * Cb + 1 for t in range(T)])
        assert np.array_equal(p.col_upper[second], BASE.col_upper[second] * factor)
    assert np.array_equal(p.A.csc_values(), BASE.A.csc_values())


def test_sour_and_diesel_regimes():
    p = scen(regime="sour_crude")
    r = p.c[COL == 0] / BASE.c[COL == 0]
    pos = np.tile(np.arange(NC), T)
    assert np.allclose(r[pos == 0], 0.85, rtol=1e-15) and np.allclose(r[pos == 1], 1.25, rtol=1e-15)
    assert np.allclose(r[pos >= 2], 1.10, rtol=1e-15)
    d = scen(regime="diesel_surge")
    sp = np.tile(np.arange(P), T)
    rs = d.c[COL == 1] / BASE.c[COL == 1]
    assert np.allclose(rs[sp % 4 == 0], 1.30 * 1.25, rtol=1e-15) and np.allclose(rs[sp % 4 != 0], 1.30, rtol=1e-15)
    ru = d.col_upper[COL == 1] / BASE.col_upper[COL == 1]
    assert np.allclose(ru[sp % 4 == 0], 1.60 / 1.25, rtol=1e-15) and np.allclose(ru[sp % 4 != 0], 1.0)


def test_multipliers_stay_in_the_reference_clip_ranges():
    p = scen(price_std=2.0, supply_std=2.0, capacity_std=2.0, demand_std=2.0)  # large std: clips bind
    crude = p.c[COL == 0] / BASE.c[COL == 0]
    sale = p.c[COL == 1] / BASE.c[COL == 1]
    supply = p.col_upper[COL == 0] / BASE.col_upper[COL == 0]
    demand = p.col_upper[COL == 1] / BASE.col_upper[COL == 1]
    cap = p.row_upper[ROW == 2] / BASE.row_upper[ROW == 2]
    for r, lo, hi in ((crude, 0.70, 1.40), (sale, 0.75, 1.35), (supply, 0.60, 1.50), (demand, 0.80, 1.25),
                      (cap, 0.75, 1.30)):
        assert np.all(r >= lo - 1e-12) and np.all(r <= hi + 1e-12)
        assert np.isclose(r.min(), lo) and np.isclose(r.max(), hi)  # both clips reached
    assert np.array_equal(p.c[COL == 2], BASE.c[COL == 2])  # inventory costs untouched
    assert np.array_equal(p.col_upper[COL == 2], BASE.col_upper[COL == 2])


def test_outage_and_random_family():
    p = scen(zero_capacity=True)
    z = np.flatnonzero(p.row_upper != BASE.row_upper)
    assert len(z) == 1 and ROW[z[0]] == 2 and p.row_upper[z[0]] == 0.0
    base = S.gen_random(50, 120, 4, 9)
    r = S.gen_scenario_ex("random", (50, 120, 4), 9, 3, shock_sigma=0.3, price_std=0.5, demand_std=0.5)
    assert np.array_equal(r.col_upper, base.col_upper) and np.array_equal(r.row_upper, base.row_upper)
    assert not np.array_equal(r.c, base.c)
    with pytest.raises(RuntimeError, match="negative"):
        scen(price_std=-1.0)
