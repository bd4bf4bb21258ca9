"""ctypes loader of the oracle (TEST INFRASTRUCTURE ONLY).

The CPU restatement of the reference GPU PDHG path lives in
pdhg_oracle.c / pdhg_ref_body.h / pdhg_enh.c and is built into
oracle/liboracle.so by oracle/Makefile.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


class RefCfg(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("tol_primal", C.c_double),
                ("tol_dual", C.c_double), ("check_frequency", C.c_int),
                ("step_tau", C.c_double), ("step_sigma", C.c_double),
                ("enable_restart", C.c_int), ("x0", C.POINTER(C.c_double)),
                ("y0", C.POINTER(C.c_double))]


class RefOut(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("checks", C.c_int),
                ("last_primal", C.c_double), ("last_dual", C.c_double),
                ("best_residual", C.c_double), ("objective", C.c_double),
                ("max_abs_c", C.c_double), ("max_abs_b", C.c_double)]


class EnhCfg(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("check_frequency", C.c_int),
                ("tol", C.c_double), ("reflection", C.c_double), ("ruiz_iters", C.c_int),
                ("threads", C.c_int), ("step", C.c_double), ("max_refinements", C.c_int),
                ("refine_ratio", C.c_double), ("x_start", C.POINTER(C.c_double)),
                ("y_start", C.POINTER(C.c_double)), ("omega_start", C.c_double)]


class EnhOut(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("restarts", C.c_int),
                ("refinements", C.c_int), ("eps_p", C.c_double), ("eps_d", C.c_double), ("eps_gap", C.c_double),
                ("kkt", C.c_double), ("primal_obj", C.c_double), ("dual_obj", C.c_double),
                ("step", C.c_double), ("omega", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        _lib = C.CDLL(str(LIB))
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class LpArrays:
    """int32 CSC arrays of a sih_solver.Problem (or raw arrays) for the oracle."""

    def __init__(self, m, n, col_ptr, row_ind, vals, c, l, u, rl, ru):
        self.m, self.n = int(m), int(n)
        self.cp = np.ascontiguousarray(col_ptr, np.int32)
        self.ri = np.ascontiguousarray(row_ind, np.int32)
        self.v = np.ascontiguousarray(vals, np.float64)
        self.c = np.ascontiguousarray(c, np.float64)
        self.l = np.ascontiguousarray(l, np.float64)
        self.u = np.ascontiguousarray(u, np.float64)
        self.rl = np.ascontiguousarray(rl, np.float64)
        self.ru = np.ascontiguousarray(ru, np.float64)

    @staticmethod
    def of(problem):
        A = problem.A
        return LpArrays(problem.num_rows(), problem.num_cols(), A.csc_col_ptr(),
                        A.csc_row_ind(), A.csc_values(), problem.c, problem.col_lower,
                        problem.col_upper, problem.row_lower, problem.row_upper)

    def args(self):
        return (self.m, self.n, _p(self.cp, C.c_int), _p(self.ri, C.c_int),
                _p(self.v, C.c_double), _p(self.c, C.c_double), _p(self.l, C.c_double),
                _p(self.u, C.c_double), _p(self.rl, C.c_double), _p(self.ru, C.c_double))


def pdhg_ref(lp: LpArrays, max_iterations=4000, tol_primal=1e-4, tol_dual=1e-4,
             check_frequency=25, step_tau=0.0, step_sigma=0.0, enable_restart=True,
             fp64=False, x0=None, y0=None):
    """Reference-mode PDHG on the CPU (gpu_solver.cpp:60-307).  Returns a dict.
    x0 / y0: an optional start point (PDHG dual sign) in place of (clamp(0), 0)."""
    cfg = RefCfg(max_iterations, tol_primal, tol_dual, check_frequency, step_tau, step_sigma,
                 int(enable_restart))
    keep = []
    for name, v in (("x0", x0), ("y0", y0)):
        if v is not None:
            a = np.ascontiguousarray(v, np.float64)
            keep.append(a)
            setattr(cfg, name, _p(a, C.c_double))
    out = RefOut()
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    dt = np.float64 if fp64 else np.float32
    ct = C.c_double if fp64 else C.c_float
    xl = np.zeros(lp.n, dt)
    xb = np.zeros(lp.n, dt)
    fn = lib().orc_pdhg_ref_f64 if fp64 else lib().orc_pdhg_ref_f32
    rc = fn(*lp.args(), C.byref(cfg), _p(x, C.c_double), _p(y, C.c_double), _p(xl, ct),
            _p(xb, ct), C.byref(out))
    assert rc == 0
    return dict(status=out.status, iterations=out.iterations, checks=out.checks,
                last_primal=out.last_primal, last_dual=out.last_dual,
                best_residual=out.best_residual, objective=out.objective,
                max_abs_c=out.max_abs_c, max_abs_b=out.max_abs_b, x=x, y=y, x_last=xl,
                xbar_last=xb)


def kkt(lp: LpArrays, x, y):
    """FP64 relative KKT in the reference dual sign; returns a dict."""
    out = np.zeros(10)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    lib().orc_kkt(*lp.args(), _p(x, C.c_double), _p(y, C.c_double), _p(out, C.c_double))
    keys = ("eps_p", "eps_d", "eps_gap", "kkt", "pobj", "dobj", "row_viol", "col_viol",
            "dual_viol", "bnorm")
    return dict(zip(keys, out.tolist()))


def canon_sum(v, is_max=False):
    """The enhanced mode's fixed-order FP64 total (orc_canon_sum)."""
    v = np.ascontiguousarray(v, np.float64)
    f = lib().orc_canon_sum
    f.restype = C.c_double
    f.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int]
    return f(_p(v, C.c_double), len(v), int(is_max))


def kkt_engine(lp: LpArrays, x, y):
    """The enhanced mode's FP64 check in the engine's order (orc_kkt_engine)."""
    out = np.zeros(6)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    lib().orc_kkt_engine(*lp.args(), _p(x, C.c_double), _p(y, C.c_double), _p(out, C.c_double))
    return dict(zip(("eps_p", "eps_d", "eps_gap", "kkt", "pobj", "dobj"), out.tolist()))


def identify_basis(lp: LpArrays, x):
    cb = np.zeros(lp.n, np.int8)
    rb = np.zeros(lp.m, np.int8)
    x = np.ascontiguousarray(x, np.float64)
    args = lp.args()
    lib().orc_identify_basis(args[0], args[1], args[2], args[3], args[4], args[6], args[7],
                             args[8], args[9], _p(x, C.c_double), _p(cb, C.c_int8),
                             _p(rb, C.c_int8))
    return cb, rb


def csc_to_csr(lp: LpArrays):
    rp = np.zeros(lp.m + 1, np.int32)
    ci = np.zeros(lp.ri.size, np.int32)
    cv = np.zeros(lp.ri.size, np.float64)
    rc = lib().orc_csc_to_csr(lp.m, lp.n, _p(lp.cp, C.c_int), _p(lp.ri, C.c_int),
                              _p(lp.v, C.c_double), _p(rp, C.c_int), _p(ci, C.c_int),
                              _p(cv, C.c_double))
    assert rc == 0
    return rp, ci, cv


def spmv_csr(m, rp, ci, v, x):
    y = np.zeros(m)
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    lib().orc_spmv_csr_f64(m, _p(rp, C.c_int), _p(ci, C.c_int), _p(v, C.c_double),
                           _p(x, C.c_double), _p(y, C.c_double))
    return y


def spmv_csc(n, cp, ri, v, y):
    z = np.zeros(n)
    cp = np.ascontiguousarray(cp, np.int32)
    ri = np.ascontiguousarray(ri, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    lib().orc_spmv_csc_f64(n, _p(cp, C.c_int), _p(ri, C.c_int), _p(v, C.c_double),
                           _p(y, C.c_double), _p(z, C.c_double))
    return z


def spmv_seq_f32(rows, ptr, idx, v, x):
    out = np.zeros(rows, np.float32)
    ptr = np.ascontiguousarray(ptr, np.int32)
    idx = np.ascontiguousarray(idx, np.int32)
    v = np.ascontiguousarray(v, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    lib().orc_spmv_seq_f32(rows, _p(ptr, C.c_int), _p(idx, C.c_int), _p(v, C.c_float),
                           _p(x, C.c_float), _p(out, C.c_float))
    return out


def spmv_enh_f32(rows, ptr, idx, v, x):
    """The enhanced mode's FP32 row sums (long rows in segments and blocks)."""
    out = np.zeros(rows, np.float32)
    ptr = np.ascontiguousarray(ptr, np.int32)
    idx = np.ascontiguousarray(idx, np.int32)
    v = np.ascontiguousarray(v, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    lib().orc_spmv_enh_f32(rows, _p(ptr, C.c_int), _p(idx, C.c_int), _p(v, C.c_float),
                           _p(x, C.c_float), _p(out, C.c_float))
    return out


def pdhg_enh(lp: LpArrays, tol=1e-6, max_iterations=100000, check_frequency=64,
             reflection=1.0, ruiz_iters=10, threads=1, step=0.0, max_refinements=8,
             refine_ratio=1e-2, x0=None, y0=None, omega0=0.0):
    """CPU port of the enhanced mode; returns a dict with the best point.
    x0 / y0 / omega0 / step: a warm start (y0 in the reference dual sign)."""
    xs = None if x0 is None else np.ascontiguousarray(x0, np.float64)
    ys = None if y0 is None else np.ascontiguousarray(y0, np.float64)
    cfg = EnhCfg(max_iterations, check_frequency, tol, reflection, ruiz_iters, threads, step,
                 max_refinements, refine_ratio, None if xs is None else _p(xs, C.c_double),
                 None if ys is None else _p(ys, C.c_double), omega0)
    out = EnhOut()
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    rc = lib().orc_pdhg_enh(*lp.args(), C.byref(cfg), _p(x, C.c_double), _p(y, C.c_double),
                            C.byref(out))
    assert rc == 0
    return dict(status=out.status, iterations=out.iterations, restarts=out.restarts,
                refinements=out.refinements, eps_p=out.eps_p, eps_d=out.eps_d, eps_gap=out.eps_gap, kkt=out.kkt,
                primal_obj=out.primal_obj, dual_obj=out.dual_obj, step=out.step, omega=out.omega, x=x, y=y)
