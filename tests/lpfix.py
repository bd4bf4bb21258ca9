"""Small LPs with known answers, shared by the CPU and GPU tests."""
import math

import numpy as np

import sih_solver as S


def toy01():
    """data/toy/toy01_lp_simple.mps: min -x1 - 2x2, x1+x2<=4, x1+3x2<=9, 0<=x<=10.
    Optimum x*=(1.5, 2.5), objective -6.5 (data/toy/toy_suite_metadata.json:5-12)."""
    p = S.Problem("toy01_lp_simple")
    p.resize(2, 2)
    p.A = S.SparseMatrix.from_triplets(
        2, 2, [S.Triplet(0, 0, 1.0), S.Triplet(1, 0, 1.0), S.Triplet(0, 1, 1.0),
               S.Triplet(1, 1, 3.0)])
    p.c = np.array([-1.0, -2.0])
    p.row_upper = np.array([4.0, 9.0])
    p.col_upper = np.array([10.0, 10.0])
    return p


def kat3x3():
    """tests/unit/test_sparse_matrix.cpp: rows [1 0 2; 0 3 0; 4 5 6]."""
    trip = [S.Triplet(0, 0, 1.0), S.Triplet(0, 2, 2.0), S.Triplet(1, 1, 3.0),
            S.Triplet(2, 0, 4.0), S.Triplet(2, 1, 5.0), S.Triplet(2, 2, 6.0)]
    return S.SparseMatrix.from_triplets(3, 3, trip)


def random_lp(m, n, density, seed, kinds=("eq", "le", "ge", "range", "free"),
              inf_bounds=True):
    """Random feasible LP with every row/bound type the reference handles."""
    rng = np.random.default_rng(seed)
    dense = rng.uniform(-1, 1, (m, n)) * (rng.random((m, n)) < density)
    for j in range(n):  # no empty column
        if not dense[:, j].any():
            dense[rng.integers(m), j] = rng.uniform(0.5, 1.0)
    x0 = rng.uniform(0, 5, n)
    ax = dense @ x0
    rl = np.empty(m)
    ru = np.empty(m)
    for i in range(m):
        k = kinds[rng.integers(len(kinds))]
        if k == "eq":
            rl[i] = ru[i] = ax[i]
        elif k == "le":
            rl[i], ru[i] = -math.inf, ax[i] + rng.uniform(0, 1)
        elif k == "ge":
            rl[i], ru[i] = ax[i] - rng.uniform(0, 1), math.inf
        elif k == "range":
            rl[i], ru[i] = ax[i] - rng.uniform(0, 1), ax[i] + rng.uniform(0, 1)
        else:
            rl[i], ru[i] = -math.inf, math.inf
    l = np.zeros(n)
    u = np.full(n, 10.0)
    if inf_bounds:
        u[rng.random(n) < 0.2] = math.inf
    cols, rows, vals = [], [], []
    for j in range(n):
        nz = np.nonzero(dense[:, j])[0]
        rows.extend(nz.tolist())
        vals.extend(dense[nz, j].tolist())
        cols.append(len(rows))
    p = S.Problem(f"rand{m}x{n}s{seed}")
    p.resize(m, n)
    p.A = S.SparseMatrix.from_csc(m, n, np.array([0] + cols), np.array(rows, np.int64),
                                  np.array(vals))
    p.c = rng.uniform(-1, 1, n) + np.abs(dense).sum(axis=0) * 0.1
    p.row_lower, p.row_upper, p.col_lower, p.col_upper = rl, ru, l, u
    return p, dense
