"""Pins the oracle (CPU restatement of the reference GPU PDHG path) against
closed forms, the reference's own known-answer tests and independent solvers."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

import oracle as O
import sih_solver as S
from lpfix import kat3x3, random_lp, toy01


def test_spmv_kat_reference_unit_test():
    # tests/unit/test_sparse_matrix.cpp:45-49 and :59-63
    A = kat3x3()
    cp, ri, v = A.csc_col_ptr(), A.csc_row_ind(), A.csc_values()
    lp = O.LpArrays(3, 3, cp, ri, v, np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3),
                    np.zeros(3))
    rp, ci, cv = O.csc_to_csr(lp)
    assert np.array_equal(O.spmv_csr(3, rp, ci, cv, [1.0, 2.0, 3.0]), [7.0, 6.0, 32.0])
    assert np.array_equal(O.spmv_csc(3, cp, ri, v, [1.0, 0.0, 1.0]), [5.0, 5.0, 8.0])
    assert np.array_equal(O.spmv_seq_f32(3, rp, ci, cv, [1, 2, 3]), [7, 6, 32])
    assert np.array_equal(O.spmv_seq_f32(3, cp, ri, v, [1, 0, 1]), [5, 5, 8])


def test_reference_csc_upload_defect_kat():
    """The reference device reads int64 CSC arrays as int32 words
    (gpu_solver.cpp:202-203); the restated A^T y on that view gives [0,5,0]
    where the correct product is [5,5,8].  The engine narrows element-wise."""
    A = kat3x3()
    cp64, ri64 = A.csc_col_ptr().astype(np.int64), A.csc_row_ind().astype(np.int64)
    d_cp = cp64.view(np.int32)[:4]  # (n+1)*sizeof(int) bytes of the int64 array
    d_ri = ri64.view(np.int32)[:6]
    assert d_cp.tolist() == [0, 0, 2, 0] and d_ri.tolist() == [0, 0, 2, 0, 1, 0]
    y = np.array([1, 0, 1], np.float32)
    out = np.zeros(3, np.float32)
    vals = A.csc_values().astype(np.float32)
    for j in range(3):  # cuda_pdhg.cu:33-49 on the corrupted arrays
        s = np.float32(0)
        for k in range(d_cp[j], d_cp[j + 1]):
            s = np.float32(s + vals[k] * y[d_ri[k]])
        out[j] = s
    assert out.tolist() == [0.0, 5.0, 0.0]
    assert O.spmv_seq_f32(3, A.csc_col_ptr(), A.csc_row_ind(), vals, y).tolist() == [5, 5, 8]


@pytest.mark.parametrize("seed", range(4))
def test_csc_to_csr_matches_dense(seed):
    p, dense = random_lp(37, 53, 0.15, seed)
    lp = O.LpArrays.of(p)
    rp, ci, cv = O.csc_to_csr(lp)
    rebuilt = np.zeros_like(dense)
    for i in range(lp.m):
        cols = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(cols) > 0)  # ascending columns within a row
        rebuilt[i, cols] = cv[rp[i]:rp[i + 1]]
    assert np.array_equal(rebuilt, dense)
    x = np.random.default_rng(seed).normal(size=lp.n)
    y = np.random.default_rng(seed + 9).normal(size=lp.m)
    np.testing.assert_allclose(O.spmv_csr(lp.m, rp, ci, cv, x), dense @ x, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.spmv_csc(lp.n, lp.cp, lp.ri, lp.v, y), dense.T @ y, rtol=1e-12,
                               atol=1e-12)


def test_first_iterate_closed_form():
    """One reference iteration from x0 = clamp(0,l,u), y0 = 0 (gpu_solver.cpp:192-196):
    x1 = clamp(x0 - tau*c), xbar = 2 x1 - x0, y1 = sigma * (A xbar - clamp(A xbar, rl, ru)),
    with tau_j = 0.95/max(1, |A_:j|_1), sigma_i = 0.95/max(1, |A_i:|_1)."""
    p, dense = random_lp(20, 30, 0.3, 7)
    lp = O.LpArrays.of(p)
    r = O.pdhg_ref(lp, max_iterations=1, check_frequency=1, fp64=True)
    tau = 0.95 / np.maximum(1.0, np.abs(dense).sum(axis=0))
    sigma = 0.95 / np.maximum(1.0, np.abs(dense).sum(axis=1))
    x0 = np.clip(0.0, lp.l, lp.u)
    x1 = np.clip(x0 - tau * lp.c, lp.l, lp.u)
    xbar = 2 * x1 - x0
    w = dense @ xbar
    y1 = sigma * (w - np.clip(w, lp.rl, lp.ru))
    np.testing.assert_allclose(r["x_last"], x1, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(r["xbar_last"], xbar, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(r["y"], y1, rtol=1e-12, atol=1e-14)
    assert r["iterations"] == 2 and r["status"] == 4  # K+1 when not converged


def _f32(q):
    """The float32 nearest the rational q (ties to even): one correct rounding."""
    from fractions import Fraction

    q = Fraction(q)
    if q == 0:
        return np.float32(0.0)
    sgn, q = (-1 if q < 0 else 1), abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    if Fraction(2) ** e > q:
        e -= 1
    quantum = Fraction(2) ** (max(e, -126) - 23)
    n, r = divmod(q, quantum)
    n = int(n)
    if r > quantum / 2 or (r == quantum / 2 and n % 2 == 1):
        n += 1
    return np.float32(float(sgn * n * quantum))


def _fmaf(a, b, c):
    """fmaf(a, b, c): a*b + c with one rounding to float32."""
    from fractions import Fraction

    return _f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def test_fp32_first_iterate_closed_form_with_contraction():
    """One FP32 reference iteration (cuda_pdhg.cu:14-103) from a random start
    point, written out with every operation rounded to float32 and nvcc's
    default contraction (-fmad=true) of `sum += v * x` in both SpMVs and of
    `x_old - step * grad` into single-rounding fmaf (emulated exactly with
    rationals); `2 x_new - x_old` is exact with or without contraction, and the
    dual step has no contractible product.  The step sizes are the host's
    sequential float sums (gpu_solver.cpp:128-151).  The FP32 oracle must match
    bit for bit -- a dropped or misplaced fma changes the last bits."""
    f32 = np.float32
    p, dense = random_lp(12, 18, 0.35, 17)
    lp = O.LpArrays.of(p)
    rng = np.random.default_rng(29)
    xs, ys = rng.uniform(-1, 11, 18), rng.normal(size=12)
    r = O.pdhg_ref(lp, max_iterations=1, check_frequency=1, x0=xs, y0=ys)
    m, n = 12, 18
    cp, ri = lp.cp, lp.ri
    v = lp.v.astype(np.float32)
    rp, ci, _ = O.csc_to_csr(lp)
    csr_v = np.array([f32(dense[i, ci[k]]) for i in range(m) for k in range(rp[i], rp[i + 1])], np.float32)
    tau, sigma = np.zeros(n, np.float32), np.zeros(m, np.float32)
    for j in range(n):
        s = f32(0)
        for k in range(cp[j], cp[j + 1]):
            s = f32(s + abs(v[k]))
        tau[j] = f32(f32(0.95) / max(f32(1), s))
    for i in range(m):
        s = f32(0)
        for k in range(rp[i], rp[i + 1]):
            s = f32(s + abs(csr_v[k]))
        sigma[i] = f32(f32(0.95) / max(f32(1), s))
    l, u = lp.l.astype(np.float32), lp.u.astype(np.float32)
    rl, ru, c = lp.rl.astype(np.float32), lp.ru.astype(np.float32), lp.c.astype(np.float32)
    x0 = np.array([max(l[j], min(u[j], f32(xs[j]))) for j in range(n)], np.float32)
    y0 = ys.astype(np.float32)
    x1, xbar = np.zeros(n, np.float32), np.zeros(n, np.float32)
    for j in range(n):
        g = f32(0)
        for k in range(cp[j], cp[j + 1]):
            g = _fmaf(v[k], y0[ri[k]], g)
        grad = f32(c[j] + g)
        xn = _fmaf(-tau[j], grad, x0[j])
        xn = l[j] if xn < l[j] else xn
        xn = u[j] if xn > u[j] else xn
        x1[j] = xn
        xbar[j] = f32(f32(2) * xn - x0[j])
    y1 = np.zeros(m, np.float32)
    for i in range(m):
        h = f32(0)
        for k in range(rp[i], rp[i + 1]):
            h = _fmaf(csr_v[k], xbar[ci[k]], h)
        w = f32(h + f32(y0[i] / sigma[i]))
        q = rl[i] if w < rl[i] else w
        q = ru[i] if q > ru[i] else q
        y1[i] = f32(sigma[i] * f32(w - q))
    assert np.array_equal(r["x_last"], x1)
    assert np.array_equal(r["xbar_last"], xbar)
    assert np.array_equal(r["y"], y1.astype(np.float64))
    # the same arithmetic without contraction differs on such data: the test can tell
    assert any(_fmaf(a, b, cc) != f32(f32(a * b) + cc)
               for a, b, cc in zip(v, y0[ri], rng.normal(size=v.size).astype(np.float32)))


def test_first_iterate_closed_form_from_a_start_point():
    """The same iteration from a given start (x0 clamped to the bounds, y0):
    x1 = clamp(x0 - tau*(c + A^T y0)), xbar = 2 x1 - x0, w = A xbar + y0/sigma,
    y1 = sigma * (w - clamp(w, rl, ru)).  A start at (clamp(0), 0) is the cold
    solve exactly."""
    p, dense = random_lp(20, 30, 0.3, 7)
    lp = O.LpArrays.of(p)
    rng = np.random.default_rng(3)
    xs, ys = rng.uniform(-2, 12, 30), rng.normal(size=20)
    r = O.pdhg_ref(lp, max_iterations=1, check_frequency=1, fp64=True, x0=xs, y0=ys)
    tau = 0.95 / np.maximum(1.0, np.abs(dense).sum(axis=0))
    sigma = 0.95 / np.maximum(1.0, np.abs(dense).sum(axis=1))
    x0 = np.clip(xs, lp.l, lp.u)
    x1 = np.clip(x0 - tau * (lp.c + dense.T @ ys), lp.l, lp.u)
    xbar = 2 * x1 - x0
    w = dense @ xbar + ys / sigma
    y1 = sigma * (w - np.clip(w, lp.rl, lp.ru))
    np.testing.assert_allclose(r["x_last"], x1, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(r["xbar_last"], xbar, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(r["y"], y1, rtol=1e-12, atol=1e-13)
    cold = O.pdhg_ref(lp, max_iterations=40, check_frequency=7)
    same = O.pdhg_ref(lp, max_iterations=40, check_frequency=7, x0=np.zeros(30), y0=np.zeros(20))
    assert np.array_equal(cold["x"], same["x"]) and np.array_equal(cold["y"], same["y"])


def test_toy01_reference_pdhg_reaches_known_optimum():
    lp = O.LpArrays.of(toy01())
    for fp64 in (False, True):
        r = O.pdhg_ref(lp, max_iterations=20000, tol_primal=1e-6, tol_dual=1e-6, fp64=fp64)
        assert r["status"] == 0
        np.testing.assert_allclose(r["x"], [1.5, 2.5], atol=2e-4)
        assert abs(r["objective"] + 6.5) < 1e-3
        np.testing.assert_allclose(r["y"], [0.5, 0.5], atol=2e-3)  # PDHG sign: c + A^T y = 0


def test_restart_kernel_is_a_no_op():
    """gpu_solver.cpp:237-239 sets xbar = x every 250 iterations; the next
    primal update overwrites xbar before any read, so iterates are identical."""
    p, _ = random_lp(40, 70, 0.2, 3)
    lp = O.LpArrays.of(p)
    a = O.pdhg_ref(lp, max_iterations=777, tol_primal=0, tol_dual=0, enable_restart=True)
    b = O.pdhg_ref(lp, max_iterations=777, tol_primal=0, tol_dual=0, enable_restart=False)
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])
    assert np.array_equal(a["x_last"], b["x_last"])


def test_iteration_limit_counter_quirk():
    p, _ = random_lp(10, 15, 0.4, 1)
    r = O.pdhg_ref(O.LpArrays.of(p), max_iterations=60, tol_primal=0, tol_dual=0,
                   check_frequency=25)
    assert r["status"] == 4 and r["iterations"] == 61 and r["checks"] == 3  # 25, 50, 60


def test_kkt_closed_forms_toy01():
    lp = O.LpArrays.of(toy01())
    k = O.kkt(lp, [1.5, 2.5], [0.5, 0.5])
    assert k["kkt"] < 1e-15 and k["pobj"] == -6.5 and k["dobj"] == -6.5
    k0 = O.kkt(lp, [0.0, 0.0], [0.0, 0.0])
    # lambda = c = (-1,-2) with finite boxes: dual objective = -10*1 - 10*2 = -30
    assert k0["eps_p"] == 0 and k0["eps_d"] == 0 and k0["dobj"] == -30.0
    assert math.isclose(k0["eps_gap"], 30.0 / 31.0, rel_tol=1e-15)
    kv = O.kkt(lp, [2.0, 2.5], [0.5, 0.5])  # row 1: 4.5 > 4, row 2: 9.5 > 9
    assert math.isclose(kv["eps_p"], 0.5 / 10.0, rel_tol=1e-15)


def _highs(lp):
    A = sp.csc_matrix((lp.v, lp.ri, lp.cp), shape=(lp.m, lp.n))
    eq = lp.rl == lp.ru
    ub_rows, ub_rhs = [], []
    fin_u = (~eq) & np.isfinite(lp.ru)
    fin_l = (~eq) & np.isfinite(lp.rl)
    Aub = sp.vstack([A[fin_u], -A[fin_l]]).tocsr()
    bub = np.concatenate([lp.ru[fin_u], -lp.rl[fin_l]])
    s = linprog(lp.c, A_ub=Aub if Aub.shape[0] else None, b_ub=bub if bub.size else None,
                A_eq=A[eq] if eq.any() else None, b_eq=lp.rl[eq] if eq.any() else None,
                bounds=list(zip(lp.l, np.where(np.isfinite(lp.u), lp.u, None))), method="highs")
    assert s.status == 0
    y = np.zeros(lp.m)  # reference PDHG sign: y = -d obj / d b
    if eq.any():
        y[eq] = -s.eqlin.marginals
    mu = -s.ineqlin.marginals if bub.size else np.zeros(0)
    nu_ = int(fin_u.sum())
    y[fin_u] += mu[:nu_]
    y[fin_l] -= mu[nu_:]
    return s, y


@pytest.mark.parametrize("seed", range(3))
def test_kkt_zero_at_highs_optimum_and_pdhg_converges_to_it(seed):
    p, _ = random_lp(30, 45, 0.25, seed)
    lp = O.LpArrays.of(p)
    s, y = _highs(lp)
    k = O.kkt(lp, s.x, y)
    assert k["kkt"] < 1e-8
    k_bad = O.kkt(lp, s.x, -y)
    assert k_bad["kkt"] > 1e-3 or np.allclose(y, 0)
    r = O.pdhg_ref(lp, max_iterations=200000, tol_primal=1e-7, tol_dual=1e-7, fp64=True)
    if r["status"] == 0:  # the reference method is slow; when it stops, it is right
        assert abs(r["objective"] - s.fun) <= 1e-3 * (1 + abs(s.fun))


def test_identify_basis_hand_computed_toy01():
    lp = O.LpArrays.of(toy01())
    cb, rb = O.identify_basis(lp, [1.5, 2.5])
    # distances: x1 1.5, x2 2.5 (interior); rows tight at upper bounds (0)
    assert cb.tolist() == [0, 0] and rb.tolist() == [2, 2]
    cb, rb = O.identify_basis(lp, [0.0, 3.0])  # row2 tight, row1 slack 1, x1 at lower
    # dists: x1 0 (lower), x2 3, row1 1, row2 0 (upper) -> basic: x2, row1
    assert cb.tolist() == [1, 0] and rb.tolist() == [0, 2]


def test_identify_basis_ties_keep_index_order():
    p = S.Problem("ties")
    p.resize(1, 3)
    p.A = S.SparseMatrix.from_triplets(1, 3, [S.Triplet(0, j, 1.0) for j in range(3)])
    p.row_lower, p.row_upper = np.array([-np.inf]), np.array([np.inf])
    p.col_upper = np.full(3, 2.0)
    cb, rb = O.identify_basis(O.LpArrays.of(p), [1.0, 1.0, 1.0])
    # row activity 3 with infinite bounds -> distance 1e30 -> basic; all columns tie at 1
    assert rb.tolist() == [0] and cb.tolist() == [1, 1, 1]


def test_enhanced_row_sums_are_sequential_up_to_8192_nonzeros():
    # Rows of at most 8192 nonzeros: one FMA chain, the same bits as the
    # sequential restatement of the reference's row sums.
    rng = np.random.default_rng(5)
    lens = np.array([0, 1, 63, 64, 65, 1000, 8191, 8192])
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = rng.integers(0, 500, ptr[-1]).astype(np.int32)
    v = rng.normal(size=ptr[-1]).astype(np.float32)
    x = rng.normal(size=500).astype(np.float32)
    a = O.spmv_enh_f32(len(lens), ptr, idx, v, x)
    b = O.spmv_seq_f32(len(lens), ptr, idx, v, x)
    assert np.array_equal(a, b)


def test_enhanced_long_rows_sum_in_segments_and_blocks():
    # Rows of more than 8192 nonzeros: segments of 64 nonzeros (FMA chains
    # from zero) added in order within blocks of 4096, blocks added in order.
    x = np.ones(1, np.float32)

    def row(v):
        return O.spmv_enh_f32(1, np.array([0, len(v)], np.int32), np.zeros(len(v), np.int32), v, x)[0]

    # 2^24 then 12,799 ones: one chain loses every +1 (2^24 + 1 rounds to
    # 2^24): 2^24.  Segment 0 is 2^24 (its 63 ones lost), the other 63 of
    # block 0 are 64 each and 2^24 + 64 k is exact: block 0 = 2^24 + 4032;
    # blocks 1 and 2 hold 4096 ones, block 3 512: 2^24 + 12,736.
    v = np.ones(12800, np.float32)
    v[0] = 2.0 ** 24
    assert row(v) == np.float32(16789952.0)
    ptr = np.array([0, 12800], np.int32)
    assert O.spmv_seq_f32(1, ptr, np.zeros(12800, np.int32), v, x)[0] == np.float32(2.0 ** 24)
    # 2^24, then a single 1 at the start of every other segment: inside block
    # 0 each +1 is lost against 2^24, so block 0 is 2^24; blocks 1 and 2 sum
    # 64 ones, block 3 eight, and those add exactly: 2^24 + 136.  One level
    # of segments added in order would keep 2^24.
    w = np.zeros(12800, np.float32)
    w[0] = 2.0 ** 24
    w[64::64] = 1.0
    assert row(w) == np.float32(16777216.0 + 136.0)
    # 8,193 ones: exact in any order
    assert row(np.ones(8193, np.float32)) == np.float32(8193.0)
    # distinct small integers sum exactly: no nonzero dropped or repeated
    u = (np.arange(12800) % 7 + 1).astype(np.float32)
    assert row(u) == np.float32(u.astype(np.int64).sum())


# ---- the enhanced mode's fixed-order FP64 reductions (CPU port) ------------
# Each case is built so that rounding alone fixes the answer: 2^53 + 1 rounds
# to 2^53 (ties to even), 2^53 + 2 is exact.  The GPU equals the port bit for
# bit (tests/test_gpu_enhanced.py), so these pin the engine's order as well.
BIG = 2.0 ** 53


def test_canon_group_butterfly_pairs_lanes_eight_apart():
    # 16-lane butterfly (o = 8, 4, 2, 1, read at lane 0): lanes 1 and 9 meet
    # first (1 + 1 = 2), and 2 then adds to 2^53 exactly.  A sequential sum
    # loses both ones; a butterfly starting at o = 1 loses the one at lane 1.
    v = np.zeros(16)
    v[0], v[1], v[9] = BIG, 1.0, 1.0
    assert O.canon_sum(v) == BIG + 2.0
    assert math.fsum(v) == BIG + 2.0 and float(np.cumsum(v)[-1]) == BIG


def test_canon_lanes_take_groups_256_apart():
    # Group partials 0 and 256 belong to lane 0 and cancel there before the
    # lanes are combined; group 1 (lane 1) then contributes its 1 exactly.  In
    # group order the 1 would be lost against 2^53 first.
    v = np.zeros(257 * 16)
    v[0], v[16], v[256 * 16] = BIG, 1.0, -BIG
    assert O.canon_sum(v) == 1.0


def test_canon_waves_are_added_in_order():
    # Lanes 0, 64 and 128 head waves 0, 1 and 2: ((2^53 + 1) + 1) + 0 = 2^53,
    # where adding waves 1 and 2 first would give 2^53 + 2.
    v = np.zeros(129 * 16)
    v[0], v[64 * 16], v[128 * 16] = BIG, 1.0, 1.0
    assert O.canon_sum(v) == BIG


def test_canon_is_exact_on_integers_and_max_is_max():
    rng = np.random.default_rng(9)
    for n in (0, 1, 15, 16, 17, 4096, 70001):
        v = rng.integers(-1000, 1000, n).astype(np.float64)
        assert O.canon_sum(v) == float(v.sum())
        a = np.abs(v)
        assert O.canon_sum(a, is_max=True) == (float(a.max()) if n else 0.0)


def test_engine_kkt_closed_forms_toy01():
    # the engine-order check gives the closed forms of the reference KKT
    lp = O.LpArrays.of(toy01())
    k = O.kkt_engine(lp, [1.5, 2.5], [0.5, 0.5])
    assert k["kkt"] < 1e-15 and k["pobj"] == -6.5 and k["dobj"] == -6.5
    k0 = O.kkt_engine(lp, [0.0, 0.0], [0.0, 0.0])
    assert k0["eps_p"] == 0 and k0["eps_d"] == 0 and k0["dobj"] == -30.0
    assert math.isclose(k0["eps_gap"], 30.0 / 31.0, rel_tol=1e-15)
    kv = O.kkt_engine(lp, [2.0, 2.5], [0.5, 0.5])
    assert math.isclose(kv["eps_p"], 0.5 / 10.0, rel_tol=1e-15)


@pytest.mark.parametrize("seed", range(3))
def test_engine_kkt_agrees_with_the_reference_kkt(seed):
    # different summation orders, the same quantities
    p, _ = random_lp(40, 70, 0.2, seed)
    lp = O.LpArrays.of(p)
    rng = np.random.default_rng(seed)
    x = np.clip(rng.normal(size=lp.n), lp.l, lp.u)
    y = rng.normal(size=lp.m)
    a, b = O.kkt_engine(lp, x, y), O.kkt(lp, x, y)
    for key in ("eps_p", "eps_d", "eps_gap", "kkt", "pobj", "dobj"):
        assert math.isclose(a[key], b[key], rel_tol=1e-12, abs_tol=1e-14), key


def test_port_warm_start_with_its_weight_refines_at_the_first_restart():
    # A start that carries its primal weight is a previous solve's point, at
    # its FP32 iterates' precision: the first restart re-centres it in FP64
    # (without that the refinement baseline is the first check's KKT, about
    # the start's own, and the solve stalls at the FP32 floor with no
    # refinement).  A start at the cold point without a weight is the cold
    # solve bit for bit; the reported weight is the effective one.
    lp = O.LpArrays.of(S.gen_staircase(4, seed=3))
    r = O.pdhg_enh(lp, tol=1e-6)
    assert r["status"] == 0 and r["omega"] > 0.0
    w = O.pdhg_enh(lp, tol=1e-7, max_iterations=60000, x0=r["x"], y0=r["y"], omega0=r["omega"], step=r["step"])
    assert w["status"] == 0 and w["refinements"] >= 1, (w["iterations"], w["kkt"])
    assert O.kkt(lp, w["x"], w["y"])["kkt"] <= 1e-7 * 1.0001
    z = O.pdhg_enh(lp, tol=1e-6, x0=np.clip(0.0, lp.l, lp.u), y0=np.zeros(lp.m))
    assert (z["iterations"], z["refinements"]) == (r["iterations"], r["refinements"])
    assert np.array_equal(z["x"], r["x"]) and np.array_equal(z["y"], r["y"]) and z["omega"] == r["omega"]
