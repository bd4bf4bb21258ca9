"""Adjoint (dL/dc, dL/db of L = w . x*) on the GPU.

Checks, per instance:
  * against the active-set closed form dL/db_R = A_RB^-T w_B (numpy lstsq on
    the active set of the returned solution), relative 1e-6;
  * against central finite differences of L along random directions of b and
    c, with FP64 HiGHS solves (relative step 1e-4), relative 1e-4 (the target);
  * against B^-1 sensitivities of the reference FP64 dual simplex's optimal
    basis (tests/golden fixtures, nondegenerate instances), relative 1e-6;
  * through torch.autograd (LpBatchSolve): the same gradients land in .grad.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import sih_solver as S
from test_golden import NAMES, fixture, problem_of
from test_oracle import _highs

pytestmark = pytest.mark.gpu


def solve(probs, tol=1e-9):
    sols = S.solve_batch(probs, S.GpuPdhgConfig.enhanced(tol=tol, max_iterations=400000))
    assert all(s.status == S.SolutionStatus.Optimal for s in sols)
    return sols


def dense(p):
    lp = O.LpArrays.of(p)
    return sp.csc_matrix((lp.v, lp.ri, lp.cp), shape=(lp.m, lp.n)).toarray()


def closed_form(p, x, y, w, tol=1e-7):
    A = dense(p)
    B = np.where((x - p.col_lower > tol * (1 + abs(x))) & (p.col_upper - x > tol * (1 + abs(x))))[0]
    # active rows: nonzero duals, or tight at x (a zero dual at a degenerate vertex)
    ax = A @ x
    fin = lambda v: np.where(np.isfinite(v), np.abs(v), 0.0)  # noqa: E731
    tight = tol * (1 + max(fin(p.row_lower).max(initial=0), fin(p.row_upper).max(initial=0)))
    at_bound = (np.isfinite(p.row_lower) & (np.abs(ax - p.row_lower) <= tight)) | \
               (np.isfinite(p.row_upper) & (np.abs(ax - p.row_upper) <= tight))
    R = np.where((np.abs(y) > tol * (1 + np.abs(y).max())) | at_bound)[0]
    mu = np.linalg.lstsq(A[np.ix_(R, B)].T, w[B], rcond=None)[0]
    g = np.zeros(p.num_rows())
    g[R] = mu
    return g, len(B), len(R)


def L_highs(p, w, db=None, dc=None):
    lp = O.LpArrays.of(p)
    if db is not None:
        lp = O.LpArrays(lp.m, lp.n, lp.cp, lp.ri, lp.v, lp.c, lp.l, lp.u, lp.rl + db, lp.ru + db)
    if dc is not None:
        lp = O.LpArrays(lp.m, lp.n, lp.cp, lp.ri, lp.v, lp.c + dc, lp.l, lp.u, lp.rl, lp.ru)
    return float(w @ _highs(lp)[0].x)


def test_adjoint_closed_form_and_finite_differences(need_gpu):
    probs = [S.gen_staircase(2, 40, 96, 7), S.gen_staircase(1, 60, 144, 8), S.gen_random(40, 96, 4, 9)]
    sols = solve(probs)
    rng = np.random.default_rng(0)
    ws = [rng.normal(size=p.num_cols()) for p in probs]
    res = S.adjoint_batch(probs, [s.x for s in sols], [s.row_duals for s in sols], ws)
    checked = 0
    for p, s, w, r in zip(probs, sols, ws, res):
        g, nb, nr = closed_form(p, s.x, s.row_duals, w)
        assert (r.basic_cols, r.active_rows) == (nb, nr), p.name
        assert np.linalg.norm(r.dl_db - g) <= 1e-6 * max(np.linalg.norm(g), 1e-12), p.name
        assert np.all(r.dl_dc == 0)
        if nb != nr:
            continue  # degenerate vertex: directional derivatives are not unique
        for _ in range(2):
            d = rng.normal(size=p.num_rows())
            d /= np.linalg.norm(d)
            h = 1e-4 * max(1.0, np.abs(np.where(np.isfinite(p.row_upper), p.row_upper, 0)).max())
            fd = (L_highs(p, w, db=h * d) - L_highs(p, w, db=-h * d)) / (2 * h)
            assert abs(fd - g @ d) <= 1e-4 * max(abs(fd), np.linalg.norm(g) * 1e-3), (p.name, fd, g @ d)
            # x* is locally constant in c only where no nonbasic reduced cost is ~0
            # (the random-equality family zeroes 40% of them: dual degenerate)
            lam = s.reduced_costs
            at_bound = (s.x - p.col_lower <= 1e-7 * (1 + abs(s.x))) | (p.col_upper - s.x <= 1e-7 * (1 + abs(s.x)))
            hc = 1e-4 * max(1.0, np.abs(p.c).max())
            if np.abs(lam[at_bound]).min(initial=np.inf) > 10 * hc * np.sqrt(p.num_cols()):
                dcv = rng.normal(size=p.num_cols())
                dcv /= np.linalg.norm(dcv)
                fdc = (L_highs(p, w, dc=hc * dcv) - L_highs(p, w, dc=-hc * dcv)) / (2 * hc)
                assert abs(fdc) <= 1e-6 * (1 + np.abs(w).sum()), (p.name, fdc)
        checked += 1
    assert checked >= 1


@pytest.mark.parametrize("name", NAMES)
def test_adjoint_matches_reference_simplex_basis(need_gpu, name):
    meta, lp, ref = fixture(name)
    basic_cols = np.where(ref["col_basis"] == 0)[0]
    active_rows = np.where(ref["row_basis"] != 0)[0]  # nonbasic slack: row at a bound
    if len(basic_cols) != len(active_rows):
        pytest.skip("not a square basis")
    A = sp.csc_matrix((lp.v, lp.ri, lp.cp), shape=(lp.m, lp.n)).toarray()
    x = ref["x"]
    # nondegenerate only: basic columns strictly inside, active rows with nonzero duals
    inside = (x[basic_cols] - lp.l[basic_cols] > 1e-7) & (lp.u[basic_cols] - x[basic_cols] > 1e-7)
    if not inside.all() or (np.abs(ref["row_duals"][active_rows]) <= 1e-9).any():
        pytest.skip("degenerate basis")
    ax = A @ x
    bn = 1 + max(np.abs(lp.rl[np.isfinite(lp.rl)]).max(initial=0), np.abs(lp.ru[np.isfinite(lp.ru)]).max(initial=0))
    slack_rows = np.setdiff1d(np.arange(lp.m), active_rows)
    if (np.minimum(np.abs(ax - lp.rl), np.abs(ax - lp.ru))[slack_rows] <= 1e-7 * bn).any():
        pytest.skip("degenerate basis (a basic slack at a bound)")
    w = np.random.default_rng(3).normal(size=lp.n)
    g = np.zeros(lp.m)
    g[active_rows] = np.linalg.solve(A[np.ix_(active_rows, basic_cols)].T, w[basic_cols])
    p = problem_of(lp)
    s = solve([p])[0]
    r = S.adjoint_batch([p], [s.x], [s.row_duals], [w])[0]
    assert np.linalg.norm(r.dl_db - g) <= 1e-6 * max(np.linalg.norm(g), 1e-12), (name, r.dl_db, g)


def test_autograd_layer_and_batching(need_gpu):
    torch = pytest.importorskip("torch")
    from sih_solver.autograd import lp_layer

    probs = [S.gen_staircase(2, 40, 96, 7 + k) for k in range(4)]
    c = torch.tensor(np.concatenate([p.c for p in probs]), dtype=torch.float64, requires_grad=True)
    b = torch.zeros(sum(p.num_rows() for p in probs), dtype=torch.float64, requires_grad=True)
    x = lp_layer(probs, c, b, S.GpuPdhgConfig.enhanced(tol=1e-9, max_iterations=400000))
    w = torch.tensor(np.random.default_rng(1).normal(size=x.numel()))
    (w * x).sum().backward()
    assert torch.all(c.grad == 0)
    sols = solve(probs)
    off = 0
    roff = 0
    for p, s in zip(probs, sols):
        n, m = p.num_cols(), p.num_rows()
        g, _, _ = closed_form(p, s.x, s.row_duals, w.numpy()[off:off + n])
        assert np.linalg.norm(b.grad.numpy()[roff:roff + m] - g) <= 1e-5 * max(np.linalg.norm(g), 1e-9)
        one = S.adjoint_batch([p], [s.x], [s.row_duals], [w.numpy()[off:off + n]])[0]
        assert np.allclose(one.dl_db, b.grad.numpy()[roff:roff + m], rtol=1e-9, atol=1e-12)
        off += n
        roff += m


@pytest.mark.parametrize("seed", [100, 104])
def test_adjoint_on_staircase_vertex_matches_finite_differences(need_gpu, seed):
    # The refinery staircase structure (4 blocks of 400 x 900), degenerate:
    # more tight rows than basic columns and zero reduced costs.  The GPU
    # adjoint is evaluated at HiGHS's optimal vertex; central differences of
    # L = w . x*(b) (FP64 HiGHS solves) along a shift that keeps the active
    # set, d = A_RB v, must agree to 1e-6 relative.
    p = S.gen_staircase(4, 400, 900, seed)
    lp = O.LpArrays.of(p)
    s, y = _highs(lp)
    x = s.x
    rng = np.random.default_rng(seed)
    w = rng.normal(size=lp.n)
    r = S.adjoint_batch([p], [x], [y], [w])[0]
    A = sp.csc_matrix((lp.v, lp.ri, lp.cp), shape=(lp.m, lp.n))
    ax = A @ x
    isb = (x - lp.l > 1e-7 * (1 + abs(x))) & (lp.u - x > 1e-7 * (1 + abs(x)))
    fin = np.concatenate([lp.rl[np.isfinite(lp.rl)], lp.ru[np.isfinite(lp.ru)]])
    tl = 1e-7 * (1 + np.abs(fin).max())
    isr = (np.abs(y) > 1e-7 * (1 + np.abs(y).max())) | (np.abs(ax - lp.rl) <= tl) | (np.abs(ax - lp.ru) <= tl)
    assert (r.basic_cols, r.active_rows) == (int(isb.sum()), int(isr.sum()))
    B, R = np.flatnonzero(isb), np.flatnonzero(isr)
    assert r.active_rows > r.basic_cols  # degenerate vertex
    v = rng.normal(size=B.size)
    d = np.zeros(lp.m)
    d[R] = A[R][:, B] @ v
    d /= np.linalg.norm(d)
    h = 1e-4 * max(1.0, np.abs(fin).max())
    Ls = []
    for sg in (1.0, -1.0):
        q = O.LpArrays(lp.m, lp.n, lp.cp, lp.ri, lp.v, lp.c, lp.l, lp.u, lp.rl + sg * h * d, lp.ru + sg * h * d)
        Ls.append(w @ _highs(q)[0].x)
    fd = (Ls[0] - Ls[1]) / (2 * h)
    assert abs(fd - r.dl_db @ d) <= 1e-6 * max(abs(fd), 1e-9), (fd, r.dl_db @ d)
