"""MI355X PDHG LP engine -- Python mirror of refinery-optimizer's GPU PDHG API.

Names, fields and defaults follow the reference's binding of its GPU path
(core/bindings/pybind_module.cpp:84-152 and :270-296,
py/sih_solver/__init__.py:81-91), so code written against
``sih_solver.solve_gpu_pdhg`` runs unchanged.  All arithmetic happens in
libmpdhg.so (HIP kernels for gfx950 behind the C ABI of include/mpdhg.h).

With no HIP device visible, ``solve_pdhg`` raises by default; the reference
runs a CPU dual simplex instead (gpu_solver.cpp:66-71), which
``GpuPdhgConfig(cpu_fallback=1)`` reproduces with the library's host FP64
simplex.  ``GpuSolver.is_cuda_available()`` tells the caller beforehand.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
from dataclasses import dataclass, field
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import _abi
from ._abi import MPDHG_MODE_ENHANCED, MPDHG_MODE_REFERENCE, MpdhgConfig, MpdhgLp, MpdhgResult

INF = math.inf
SIH_INFINITY = math.inf


class SolutionStatus(enum.IntEnum):  # solution.hpp:11-21
    Optimal = 0
    Infeasible = 1
    Unbounded = 2
    TimeLimit = 3
    IterationLimit = 4
    NodeLimit = 5
    NumericalFailure = 6
    Interrupted = 7
    Unknown = 8


class BasisStatus(enum.IntEnum):  # solution.hpp:23-29
    Basic = 0
    AtLower = 1
    AtUpper = 2
    Free = 3
    Superbasic = 4


class ObjectiveSense(enum.IntEnum):  # problem.hpp:11-14
    Minimize = 1
    Maximize = -1


class VariableType(enum.IntEnum):  # problem.hpp:19-24
    Continuous = 0
    Binary = 1
    Integer = 2
    SemiContinuous = 3


@dataclass
class Triplet:  # sparse_matrix.hpp:10-17
    row: int = 0
    col: int = 0
    value: float = 0.0


class SparseMatrix:
    """CSC + CSR matrix with int64 indices and double values (sparse_matrix.hpp:20-77)."""

    def __init__(self, rows: int = 0, cols: int = 0):
        self._m, self._n = int(rows), int(cols)
        self._cp = np.zeros(self._n + 1, np.int64)
        self._ri = np.zeros(0, np.int64)
        self._cv = np.zeros(0, np.float64)
        self._csr = None

    @staticmethod
    def from_triplets(m: int, n: int, triplets: Iterable[Triplet], zero_tol: float = 1e-15):
        """sparse_matrix.cpp:13-88: duplicates summed, |v| <= zero_tol dropped,
        CSC ordered by (col, row)."""
        acc = {}
        for t in triplets:
            r, c_ = int(t.row), int(t.col)
            if not (0 <= r < m and 0 <= c_ < n):
                raise IndexError("Triplet indices out of matrix bounds")
            acc[(c_, r)] = acc.get((c_, r), 0.0) + float(t.value)
        keys = sorted(k for k, v in acc.items() if abs(v) > zero_tol)
        cols = np.array([k[0] for k in keys], np.int64)
        rows = np.array([k[1] for k in keys], np.int64)
        vals = np.array([acc[k] for k in keys], np.float64)
        cp = np.zeros(n + 1, np.int64)
        np.add.at(cp, cols + 1, 1)
        return SparseMatrix.from_csc(m, n, np.cumsum(cp), rows, vals)

    @staticmethod
    def from_csc(m: int, n: int, col_ptr, row_ind, vals):
        A = SparseMatrix(m, n)
        A._cp = np.ascontiguousarray(col_ptr, np.int64)
        A._ri = np.ascontiguousarray(row_ind, np.int64)
        A._cv = np.ascontiguousarray(vals, np.float64)
        if A._cp.shape != (n + 1,) or A._ri.shape != A._cv.shape or A._cp[-1] != A._ri.size:
            raise ValueError("inconsistent CSC arrays")
        return A

    def num_rows(self) -> int:
        return self._m

    def num_cols(self) -> int:
        return self._n

    def num_nonzeros(self) -> int:
        return int(self._cv.size)

    def csc_col_ptr(self):
        return self._cp

    def csc_row_ind(self):
        return self._ri

    def csc_values(self):
        return self._cv

    def _build_csr(self):
        if self._csr is None:
            cols = np.repeat(np.arange(self._n, dtype=np.int64), np.diff(self._cp))
            order = np.lexsort((cols, self._ri))
            rp = np.zeros(self._m + 1, np.int64)
            np.add.at(rp, self._ri + 1, 1)
            self._csr = (np.cumsum(rp), cols[order], self._cv[order])
        return self._csr

    def csr_row_ptr(self):
        return self._build_csr()[0]

    def csr_col_ind(self):
        return self._build_csr()[1]

    def csr_values(self):
        return self._build_csr()[2]

    def mat_vec(self, x):
        x = np.asarray(x, np.float64)
        if x.size != self._n:
            raise ValueError("Vector length does not match matrix column count in mat_vec")
        rp, ci, v = self._build_csr()
        out = np.zeros(self._m)
        np.add.at(out, np.repeat(np.arange(self._m), np.diff(rp)), v * x[ci])
        return out

    def mat_trans_vec(self, y):
        y = np.asarray(y, np.float64)
        if y.size != self._m:
            raise ValueError("Vector length does not match matrix row count in mat_trans_vec")
        out = np.zeros(self._n)
        np.add.at(out, np.repeat(np.arange(self._n), np.diff(self._cp)), self._cv * y[self._ri])
        return out


class Problem:
    """The LP fields of model::Problem (problem.hpp:37-137) that the GPU path reads."""

    def __init__(self, name: str = "problem"):
        self.name = name
        self.sense = ObjectiveSense.Minimize
        self.obj_offset = 0.0
        self._A = SparseMatrix(0, 0)
        self.c = np.zeros(0)
        self.row_lower = np.zeros(0)
        self.row_upper = np.zeros(0)
        self.col_lower = np.zeros(0)
        self.col_upper = np.zeros(0)
        self.var_types = np.zeros(0, np.int8)  # VariableType per column
        self.row_names: List[str] = []
        self.col_names: List[str] = []
        self.Q: Optional[SparseMatrix] = None  # quadratic objective 1/2 x^T Q x (read, not solved)
        self.row_kind = None  # staircase workloads: 0 balance, 1 blend, 2 capacity

    def resize(self, num_rows: int, num_cols: int):  # problem.cpp:8-23
        self._A = SparseMatrix(num_rows, num_cols)
        self.c = np.zeros(num_cols)
        self.row_lower = np.full(num_rows, -INF)
        self.row_upper = np.full(num_rows, INF)
        self.col_lower = np.zeros(num_cols)
        self.col_upper = np.full(num_cols, INF)
        self.var_types = np.zeros(num_cols, np.int8)
        self.row_names, self.col_names = [], []

    @property
    def A(self) -> SparseMatrix:
        return self._A

    @A.setter
    def A(self, mat: SparseMatrix):
        if mat.num_rows() != self.num_rows() or mat.num_cols() != self.num_cols():
            raise ValueError("Constraint matrix A dimensions do not match problem dimensions")
        self._A = mat

    def num_rows(self) -> int:
        return self._A.num_rows()

    def num_cols(self) -> int:
        return self._A.num_cols()

    def num_nonzeros(self) -> int:
        return self._A.num_nonzeros()

    def num_integers(self) -> int:  # problem.cpp:93-101
        vt = np.asarray(self.var_types)
        return int(np.count_nonzero((vt == VariableType.Integer) | (vt == VariableType.Binary)))

    def is_mip(self) -> bool:  # problem.cpp:113-115; the PDHG path ignores integrality
        return self.num_integers() > 0

    def has_quadratic(self) -> bool:
        return self.Q is not None

    def is_qp(self) -> bool:  # problem.hpp:51; the PDHG path ignores Q
        return self.Q is not None and self.Q.num_nonzeros() > 0

    # -- C view ------------------------------------------------------------
    def _c_view(self):
        """An MpdhgLp pointing into this problem's arrays (plus keep-alive refs)."""
        m, n = self.num_rows(), self.num_cols()
        arrs = [
            np.ascontiguousarray(self._A.csc_col_ptr(), np.int64),
            np.ascontiguousarray(self._A.csc_row_ind(), np.int64),
            np.ascontiguousarray(self._A.csc_values(), np.float64),
            np.ascontiguousarray(self.c, np.float64),
            np.ascontiguousarray(self.col_lower, np.float64),
            np.ascontiguousarray(self.col_upper, np.float64),
            np.ascontiguousarray(self.row_lower, np.float64),
            np.ascontiguousarray(self.row_upper, np.float64),
        ]
        if any(a.size != s for a, s in zip(arrs[3:], (n, n, n, m, m))):
            raise ValueError("vector sizes do not match the problem dimensions")
        lp = MpdhgLp()
        lp.m, lp.n, lp.nnz = m, n, self._A.num_nonzeros()
        lp.col_ptr = arrs[0].ctypes.data_as(C.POINTER(C.c_int64))
        lp.row_ind = arrs[1].ctypes.data_as(C.POINTER(C.c_int64))
        lp.vals = arrs[2].ctypes.data_as(C.POINTER(C.c_double))
        for name, a in zip(("c", "col_lower", "col_upper", "row_lower", "row_upper"), arrs[3:]):
            setattr(lp, name, a.ctypes.data_as(C.POINTER(C.c_double)))
        return lp, arrs


@dataclass
class Solution:
    """model::Solution (solution.hpp:37-84) fields the GPU path fills, plus metrics."""

    status: SolutionStatus = SolutionStatus.Unknown
    x: np.ndarray = field(default_factory=lambda: np.zeros(0))
    slack: np.ndarray = field(default_factory=lambda: np.zeros(0))
    row_duals: np.ndarray = field(default_factory=lambda: np.zeros(0))
    reduced_costs: np.ndarray = field(default_factory=lambda: np.zeros(0))
    col_basis: List[BasisStatus] = field(default_factory=list)
    row_basis: List[BasisStatus] = field(default_factory=list)
    primal_objective: float = 0.0
    dual_bound: float = 0.0
    mip_gap: float = 0.0
    simplex_iterations: int = 0
    barrier_iterations: int = 0
    nodes_explored: int = 0
    time_wall_sec: float = 0.0
    time_cpu_sec: float = 0.0
    time_presolve_sec: float = 0.0
    time_solver_sec: float = 0.0
    # extension metrics
    restarts: int = 0
    refinements: int = 0
    checks: int = 0
    rel_primal: float = math.nan
    rel_dual: float = math.nan
    rel_gap: float = math.nan
    rel_kkt: float = math.nan
    ref_primal: float = math.nan
    ref_dual: float = math.nan
    kernel_ms_spmv_csr: float = 0.0
    kernel_ms_spmv_csc: float = 0.0
    kernel_launches_csr: int = 0
    kernel_launches_csc: int = 0
    bytes_csr_launch: float = 0.0
    bytes_csc_launch: float = 0.0
    step_size: float = 0.0
    primal_weight: float = 0.0
    csr_block: int = 0  # tile schedule used for A x (threads per workgroup, LDS chunk)
    csr_chunk: int = 0
    csc_block: int = 0  # and for A^T y
    csc_chunk: int = 0
    tune_ms: float = 0.0  # time of the schedule search
    csr_bands: int = 1  # column bands of the A xbar product
    vertex_status: SolutionStatus = SolutionStatus.Unknown  # vertex_check: crossover status
    vertex_objective: float = math.nan  # vertex_check: objective of the FP64 vertex
    vertex_rel_error: float = math.nan  # vertex_check: |objective - vertex| / max(1, |vertex|)

    def is_optimal(self) -> bool:
        return self.status == SolutionStatus.Optimal

    def is_feasible(self) -> bool:
        return self.status in (SolutionStatus.Optimal, SolutionStatus.TimeLimit,
                               SolutionStatus.IterationLimit, SolutionStatus.NodeLimit)


@dataclass
class GpuPdhgConfig:
    """GpuPdhgConfig (gpu_solver.hpp:12-22) plus engine extensions."""

    max_iterations: int = 4000
    tol_primal: float = 1e-4
    tol_dual: float = 1e-4
    tol_gap: float = 1e-4
    check_frequency: int = 25
    step_tau: float = 0.0
    step_sigma: float = 0.0
    enable_restart: bool = True
    enable_hybrid_polish: bool = True
    # extensions
    mode: int = MPDHG_MODE_REFERENCE
    tol_kkt: float = 0.0
    max_refinements: int = 0
    reflection: float = 1.0
    ruiz_iterations: int = 10
    profile: bool = False
    dual_sign: int = 0
    refine_ratio: float = 1e-2
    spmv_layout: int = 0  # 0: gather order where it applies, 1: storage order
    spmv_chunk: int = 0  # nonzeros per LDS pass: 0 = auto, or 1024 / 2048 / 4096 / 8192
    launch_graph: int = 1  # replay each full check cycle as a recorded HIP graph
    spmv_block: int = 0  # threads per SpMV workgroup: 0 = auto, or 128 / 256 / 512
    spmv_tune: int = 1  # 1: time the schedules on the batch when it is large, 2: always, 0: rule
    cpu_fallback: int = 0  # 1: with no HIP device, solve on the host FP64 simplex (gpu_solver.cpp:66-71)
    spmv_bands: int = 0  # one LP: column bands of A xbar (0 rule, 1 none, k bands)
    vertex_check: int = 0  # k > 0: FP64 crossover of every LP with <= k rows after the solve

    @staticmethod
    def enhanced(tol: float = 1e-6, max_iterations: int = 100000, check_frequency: int = 64,
                 max_refinements: int = 8, **kw) -> "GpuPdhgConfig":
        """Configuration of the enhanced mode (restarted Halpern PDHG, FP64 KKT target)."""
        c = GpuPdhgConfig(max_iterations=max_iterations, check_frequency=check_frequency,
                          mode=MPDHG_MODE_ENHANCED, tol_kkt=tol, max_refinements=max_refinements)
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    def _c(self) -> MpdhgConfig:
        c = MpdhgConfig()
        _abi.lib().mpdhg_config_default(C.byref(c))
        for name, _ in MpdhgConfig._fields_:
            setattr(c, name, type(getattr(c, name))(getattr(self, name)))
        return c


@dataclass
class GpuMemoryInfo:  # gpu_solver.hpp:24-28
    total_bytes: int = 0
    free_bytes: int = 0
    used_bytes: int = 0


def _result_to_solution(r: MpdhgResult, x, y, rc, ax) -> Solution:
    s = Solution()
    s.status = SolutionStatus(r.status)
    s.x, s.row_duals, s.reduced_costs, s.slack = x, y, rc, ax
    s.primal_objective = r.primal_objective
    s.dual_bound = r.dual_objective
    s.simplex_iterations = int(r.iterations)
    s.time_wall_sec = r.time_wall_sec
    s.time_solver_sec = r.solve_sec
    for name in ("restarts", "refinements", "checks", "rel_primal", "rel_dual", "rel_gap",
                 "rel_kkt", "ref_primal", "ref_dual", "kernel_ms_spmv_csr",
                 "kernel_ms_spmv_csc", "kernel_launches_csr", "kernel_launches_csc",
                 "bytes_csr_launch", "bytes_csc_launch", "step_size", "primal_weight",
                 "csr_block", "csr_chunk", "csc_block", "csc_chunk", "tune_ms", "csr_bands",
                 "vertex_objective", "vertex_rel_error"):
        setattr(s, name, getattr(r, name))
    s.vertex_status = SolutionStatus(r.vertex_status)
    return s


@dataclass
class WarmStart:
    """Where a solve starts (mpdhg_warm_start): usually the solution of a base
    LP that a scenario perturbs (py/refinery/reoptimization.py:31-79 re-solves
    each perturbation from the previous solution).  y is in the config's
    dual_sign; primal_weight / step_size <= 0 leave those to the solver."""

    x: Optional[np.ndarray] = None
    y: Optional[np.ndarray] = None
    primal_weight: float = 0.0
    step_size: float = 0.0

    @staticmethod
    def of(sol: "Solution") -> "WarmStart":
        return WarmStart(sol.x, sol.row_duals, sol.primal_weight, sol.step_size)


class _Ctx:
    """A libmpdhg context (device buffers + one HIP stream) for reuse across solves."""

    def __init__(self, device: int = 0):
        self._p = C.c_void_p()
        _abi.check(_abi.lib().mpdhg_create(device, C.byref(self._p)), "mpdhg_create")

    def close(self):
        if self._p:
            _abi.lib().mpdhg_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[_Ctx] = None


def _ctx() -> _Ctx:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = _Ctx(0)
    return _default_ctx


def _solve_ctx(config: GpuPdhgConfig):
    """The context a solve runs on; none when no device is visible and the
    configuration asks for the host fallback (the library then runs it)."""
    if config.cpu_fallback and not GpuSolver.is_cuda_available():
        return None
    return _ctx()._p


def solve_batch(problems: Sequence[Problem], config: Optional[GpuPdhgConfig] = None,
                warm: Optional[Sequence[Optional[WarmStart]]] = None) -> List[Solution]:
    """Solve many LPs in one device-resident batch (extension of the reference
    API); warm: one WarmStart (or Solution, or None) per LP."""
    config = config or GpuPdhgConfig()
    cfg = config._c()
    nb = len(problems)
    lps = (MpdhgLp * max(nb, 1))()
    res = (MpdhgResult * max(nb, 1))()
    keep, outs = [], []
    ws = None
    if warm is not None:
        if len(warm) != nb:
            raise ValueError("one warm start per problem")
        ws = (_abi.MpdhgWarmStart * max(nb, 1))()
        dp = C.POINTER(C.c_double)
        for q, w in enumerate(warm):
            if w is None:
                continue
            if isinstance(w, Solution):
                w = WarmStart.of(w)
            m, n = problems[q].num_rows(), problems[q].num_cols()
            if w.x is not None:
                xa = np.ascontiguousarray(w.x, np.float64)
                if xa.shape != (n,):
                    raise ValueError("warm x has the wrong length")
                keep.append(xa)
                ws[q].x = xa.ctypes.data_as(dp)
            if w.y is not None:
                ya = np.ascontiguousarray(w.y, np.float64)
                if ya.shape != (m,):
                    raise ValueError("warm y has the wrong length")
                keep.append(ya)
                ws[q].y = ya.ctypes.data_as(dp)
            ws[q].primal_weight = float(w.primal_weight)
            ws[q].step_size = float(w.step_size)
    for q, p in enumerate(problems):
        lp, arrs = p._c_view()
        lps[q] = lp
        keep.append(arrs)
        m, n = p.num_rows(), p.num_cols()
        o = (np.zeros(n), np.zeros(m), np.zeros(n), np.zeros(m))
        outs.append(o)
        res[q].x = o[0].ctypes.data_as(C.POINTER(C.c_double))
        res[q].y = o[1].ctypes.data_as(C.POINTER(C.c_double))
        res[q].reduced_costs = o[2].ctypes.data_as(C.POINTER(C.c_double))
        res[q].row_activity = o[3].ctypes.data_as(C.POINTER(C.c_double))
    if ws is None:
        _abi.check(_abi.lib().mpdhg_solve_batch(_solve_ctx(config), lps, nb, C.byref(cfg), res),
                   "mpdhg_solve_batch")
    else:
        _abi.check(_abi.lib().mpdhg_solve_warm(_solve_ctx(config), lps, nb, C.byref(cfg), ws, res),
                   "mpdhg_solve_warm")
    return [_result_to_solution(res[q], *outs[q]) for q in range(nb)]


class GpuSolver:
    """gpu_solver.hpp:30-45 static interface."""

    @staticmethod
    def is_cuda_available() -> bool:  # gpu_solver.cpp:42-46
        n = C.c_int(0)
        return _abi.lib().mpdhg_device_count(C.byref(n)) == 0 and n.value > 0

    @staticmethod
    def get_vram_info() -> GpuMemoryInfo:  # gpu_solver.cpp:48-58
        info = GpuMemoryInfo()
        if not GpuSolver.is_cuda_available():
            return info
        t, f, u = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _abi.check(_abi.lib().mpdhg_mem_info(0, C.byref(t), C.byref(f), C.byref(u)),
                   "mpdhg_mem_info")
        return GpuMemoryInfo(t.value, f.value, u.value)

    @staticmethod
    def solve_pdhg(problem: Problem, config: Optional[GpuPdhgConfig] = None) -> Solution:
        return solve_batch([problem], config)[0]

    @staticmethod
    def solve_hybrid(problem: Problem, config: Optional[GpuPdhgConfig] = None) -> Solution:
        """gpu_solver.cpp:309-417: PDHG on the GPU, the basis hand-off, the FP64
        simplex polish from that basis, and the slack-basis simplex when the
        polish does not end optimal.  The Solution is the polished one: simplex
        duals (reduced cost = c - A^T y), basis statuses, simplex_iterations
        counting the pivots, time_wall_sec the whole call.  The PDHG stage's
        figures are kept in the extension fields (rel_kkt, restarts, ...)."""
        config = config or GpuPdhgConfig()
        cfg = config._c()
        lp, keep = problem._c_view()
        m, n = problem.num_rows(), problem.num_cols()
        o = _abi.MpdhgPolishOpts()
        _abi.lib().mpdhg_polish_opts_default(C.byref(o))
        o.obj_offset = float(problem.obj_offset)
        pd = MpdhgResult()
        out = (np.zeros(n), np.zeros(m), np.zeros(n), np.zeros(m), np.zeros(n, np.int8), np.zeros(m, np.int8))
        r = _abi.MpdhgPolishResult()
        dp, bp = C.POINTER(C.c_double), C.POINTER(C.c_int8)
        r.x, r.y, r.reduced_costs, r.row_activity = (a.ctypes.data_as(dp) for a in out[:4])
        r.col_basis_out, r.row_basis_out = (a.ctypes.data_as(bp) for a in out[4:])
        _abi.check(_abi.lib().mpdhg_solve_hybrid(_solve_ctx(config), C.byref(lp), C.byref(cfg), C.byref(o),
                                                 C.byref(pd), C.byref(r)), "mpdhg_solve_hybrid")
        s = _result_to_solution(pd, np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0))
        s.status = SolutionStatus(r.status)
        s.x, s.row_duals, s.reduced_costs, s.slack = out[0], out[1], out[2], out[3]
        s.col_basis = [BasisStatus(int(v)) for v in out[4]]
        s.row_basis = [BasisStatus(int(v)) for v in out[5]]
        s.primal_objective = r.objective
        s.dual_bound = r.objective
        s.simplex_iterations = int(r.iterations)
        s.time_wall_sec = r.time_sec
        s.time_solver_sec = pd.solve_sec
        return s

    @staticmethod
    def solve_batch(problems: Sequence[Problem], config: Optional[GpuPdhgConfig] = None, warm=None):
        return solve_batch(problems, config, warm)


@dataclass
class AdjointResult:
    """Gradients of L = w . x* at an LP solution (extension; see mpdhg_adjoint_batch)."""

    dl_dc: np.ndarray
    dl_db: np.ndarray   # w.r.t. a shift of each row's bounds (rl and ru together)
    dl_dl: np.ndarray   # columns held at their lower bound
    dl_du: np.ndarray   # columns held at their upper bound
    basic_cols: int
    active_rows: int
    iterations: int
    residual: float


def adjoint_batch(problems: Sequence[Problem], xs, ys, ws, active_tol: float = 1e-7,
                  tol: float = 1e-10, max_iter: int = 200000) -> List[AdjointResult]:
    """Batched adjoint on the GPU: for each LP, the gradients of w . x* at the
    solution (x, y) (row duals in the reference sign, as solve_* return them).
    CGLS needs a number of steps that grows with the coupling length of the
    active set (about 900 per time block of a staircase LP); max_iter bounds it."""
    nb = len(problems)
    lps = (MpdhgLp * max(nb, 1))()
    io = (_abi.MpdhgAdjointIO * max(nb, 1))()
    keep, outs = [], []
    dp = C.POINTER(C.c_double)
    for q, p in enumerate(problems):
        lp, arrs = p._c_view()
        lps[q] = lp
        ins = [np.ascontiguousarray(a, np.float64) for a in (xs[q], ys[q], ws[q])]
        n, m = p.num_cols(), p.num_rows()
        if ins[0].size != n or ins[1].size != m or ins[2].size != n:
            raise ValueError("x, y, w sizes do not match the problem")
        o = (np.zeros(n), np.zeros(m), np.zeros(n), np.zeros(n))
        keep.append((arrs, ins))
        outs.append(o)
        io[q].x, io[q].y, io[q].w = (a.ctypes.data_as(dp) for a in ins)
        io[q].dl_dc, io[q].dl_db, io[q].dl_dl, io[q].dl_du = (a.ctypes.data_as(dp) for a in o)
    _abi.check(_abi.lib().mpdhg_adjoint_batch(_ctx()._p, lps, nb, io, active_tol, tol, max_iter),
               "mpdhg_adjoint_batch")
    return [AdjointResult(*outs[q], io[q].basic_cols, io[q].active_rows, io[q].iterations,
                          io[q].residual) for q in range(nb)]


def identify_basis(problem: Problem, x) -> tuple:
    """Basis statuses from a primal point (gpu_solver.cpp:319-395), as int8 arrays."""
    lp, keep = problem._c_view()
    x = np.ascontiguousarray(x, np.float64)
    cb = np.zeros(problem.num_cols(), np.int8)
    rb = np.zeros(problem.num_rows(), np.int8)
    _abi.check(_abi.lib().mpdhg_identify_basis(
        C.byref(lp), x.ctypes.data_as(C.POINTER(C.c_double)),
        cb.ctypes.data_as(C.POINTER(C.c_int8)), rb.ctypes.data_as(C.POINTER(C.c_int8))),
        "mpdhg_identify_basis")
    return cb, rb


@dataclass
class PolishResult:
    """Result of the FP64 crossover (SimplexSolver::solve_from_basis)."""

    status: SolutionStatus
    x: np.ndarray
    row_duals: np.ndarray      # simplex sign: reduced cost = c - A^T y
    reduced_costs: np.ndarray
    slack: np.ndarray          # A x
    col_basis: np.ndarray      # BasisStatus values (int8)
    row_basis: np.ndarray
    objective: float
    iterations: int
    warm_start: bool
    refactorizations: int
    repaired: int
    bound_flips: int
    degenerate_pivots: int
    push_steps: int
    primal_infeasibility: float
    time_sec: float


def polish(problem: Problem, col_basis=None, row_basis=None, x_start=None, max_iterations: int = -1,
           primal_tol: float = 1e-6, dual_tol: float = 1e-6, refactor_every: int = 100,
           time_limit_sec: float = 0.0) -> PolishResult:
    """FP64 bounded primal simplex from a basis (simplex_solver.cpp:266-329),
    in libmpdhg (host code).  With no basis, or one without exactly m basic
    entries, it starts from the slack basis.  x_start (the point the basis
    came from) adds a primal push of the variables it puts inside their
    bounds before the simplex."""
    lp, keep = problem._c_view()
    m, n = problem.num_rows(), problem.num_cols()
    o = _abi.MpdhgPolishOpts()
    _abi.lib().mpdhg_polish_opts_default(C.byref(o))
    o.max_iterations, o.primal_tol, o.dual_tol = int(max_iterations), float(primal_tol), float(dual_tol)
    o.obj_offset, o.refactor_every, o.time_limit_sec = float(problem.obj_offset), int(refactor_every), float(time_limit_sec)
    out = (np.zeros(n), np.zeros(m), np.zeros(n), np.zeros(m), np.zeros(n, np.int8), np.zeros(m, np.int8))
    r = _abi.MpdhgPolishResult()
    dp, bp = C.POINTER(C.c_double), C.POINTER(C.c_int8)
    r.x, r.y, r.reduced_costs, r.row_activity = (a.ctypes.data_as(dp) for a in out[:4])
    r.col_basis_out, r.row_basis_out = (a.ctypes.data_as(bp) for a in out[4:])
    cb = rb = None
    if col_basis is not None and row_basis is not None:
        cb = np.ascontiguousarray(np.asarray(col_basis, np.int8))
        rb = np.ascontiguousarray(np.asarray(row_basis, np.int8))
        if cb.size != n or rb.size != m:
            raise ValueError("basis sizes do not match the problem")
    xs = None
    if x_start is not None:
        xs = np.ascontiguousarray(np.asarray(x_start, np.float64))
        if xs.size != n:
            raise ValueError("x_start size does not match the problem")
    _abi.check(_abi.lib().mpdhg_polish(C.byref(lp), cb.ctypes.data_as(bp) if cb is not None else None,
                                       rb.ctypes.data_as(bp) if rb is not None else None,
                                       xs.ctypes.data_as(dp) if xs is not None else None,
                                       C.byref(o), C.byref(r)), "mpdhg_polish")
    return PolishResult(SolutionStatus(r.status), out[0], out[1], out[2], out[3], out[4], out[5],
                        r.objective, int(r.iterations), bool(r.warm_start), int(r.refactorizations),
                        int(r.repaired), int(r.bound_flips), int(r.degenerate_pivots),
                        int(r.push_steps), r.primal_infeasibility, r.time_sec)


def solve_gpu_pdhg(problem: Problem, config: Optional[GpuPdhgConfig] = None) -> Solution:
    """py/sih_solver/__init__.py:81-85."""
    if config is None:
        config = GpuPdhgConfig()
    return GpuSolver.solve_pdhg(problem, config)


def solve_gpu_hybrid(problem: Problem, config: Optional[GpuPdhgConfig] = None) -> Solution:
    """py/sih_solver/__init__.py:87-91."""
    if config is None:
        config = GpuPdhgConfig()
    return GpuSolver.solve_hybrid(problem, config)


# ---- synthetic workloads --------------------------------------------------
def _owned_to_problem(ptr: C.c_void_p, name: str) -> Problem:
    L = _abi.lib()
    v = L.mpdhg_lp_view(ptr).contents
    m, n, nnz = v.m, v.n, v.nnz

    def arr(p, k, dt):
        return np.ctypeslib.as_array(p, shape=(k,)).astype(dt, copy=True) if k else np.zeros(0, dt)

    p = Problem(name)
    p._A = SparseMatrix.from_csc(m, n, arr(v.col_ptr, n + 1, np.int64),
                                 arr(v.row_ind, nnz, np.int64), arr(v.vals, nnz, np.float64))
    p.c = arr(v.c, n, np.float64)
    p.col_lower = arr(v.col_lower, n, np.float64)
    p.col_upper = arr(v.col_upper, n, np.float64)
    p.row_lower = arr(v.row_lower, m, np.float64)
    p.row_upper = arr(v.row_upper, m, np.float64)
    p.var_types = np.zeros(n, np.int8)
    rk = L.mpdhg_lp_row_kind(ptr)
    if rk:
        p.row_kind = np.ctypeslib.as_array(rk, shape=(m,)).astype(np.int8, copy=True)
    L.mpdhg_lp_free(ptr)
    return p


def gen_random(m: int, n: int, nnz_per_col: int, seed: int) -> Problem:
    ptr = C.c_void_p()
    _abi.check(_abi.lib().mpdhg_gen_random(m, n, nnz_per_col, seed, C.byref(ptr)), "gen_random")
    return _owned_to_problem(ptr, f"random_{m}x{n}_s{seed}")


def gen_staircase(T: int, R: int = 400, Cb: int = 900, seed: int = 0) -> Problem:
    ptr = C.c_void_p()
    _abi.check(_abi.lib().mpdhg_gen_staircase(T, R, Cb, seed, C.byref(ptr)), "gen_staircase")
    return _owned_to_problem(ptr, f"staircase_T{T}_s{seed}")


def gen_scenario(kind: str, size: tuple, base_seed: int, scenario_seed: int,
                 shock_sigma: float = 0.3, zero_capacity: bool = False) -> Problem:
    """A price-shocked scenario of a base workload: kind "random" with size
    (m, n, nnz_per_col) or kind "staircase" with size (T, R, C)."""
    L = _abi.lib()
    base = C.c_void_p()
    if kind == "random":
        _abi.check(L.mpdhg_gen_random(*size, base_seed, C.byref(base)), "gen_random")
    elif kind == "staircase":
        _abi.check(L.mpdhg_gen_staircase(*size, base_seed, C.byref(base)), "gen_staircase")
    else:
        raise ValueError(f"unknown workload kind {kind!r}")
    out = C.c_void_p()
    rc = L.mpdhg_gen_scenario(base, scenario_seed, shock_sigma, int(zero_capacity), C.byref(out))
    L.mpdhg_lp_free(base)
    _abi.check(rc, "gen_scenario")
    return _owned_to_problem(out, f"{kind}_{'x'.join(map(str, size))}_s{base_seed}_{scenario_seed}")


REGIMES = {"normal": 0, "high_crude": 1, "diesel_surge": 2, "turnaround": 3, "sour_crude": 4,
           "sweet_crude_shortage": 5}


def gen_scenario_ex(kind: str, size: tuple, base_seed: int, scenario_seed: int, shock_sigma: float = 0.0,
                    price_std: float = 0.0, supply_std: float = 0.0, capacity_std: float = 0.0,
                    demand_std: float = 0.0, regime: str = "normal", zero_capacity: bool = False) -> Problem:
    """A perturbed scenario of a base workload (mpdhg_gen_scenario_ex): the
    perturbations of py/refinery/parameterized_generator.py:28-133 on the
    staircase model (crude prices and supplies, unit capacities, product
    prices and demand caps, market regimes); random-equality LPs take the
    lognormal price shock only."""
    L = _abi.lib()
    base = C.c_void_p()
    if kind == "random":
        _abi.check(L.mpdhg_gen_random(*size, base_seed, C.byref(base)), "gen_random")
    elif kind == "staircase":
        _abi.check(L.mpdhg_gen_staircase(*size, base_seed, C.byref(base)), "gen_staircase")
    else:
        raise ValueError(f"unknown workload kind {kind!r}")
    o = _abi.MpdhgScenarioOpts()
    L.mpdhg_scenario_opts_default(C.byref(o))
    o.shock_sigma, o.price_std, o.supply_std = shock_sigma, price_std, supply_std
    o.capacity_std, o.demand_std = capacity_std, demand_std
    o.regime = REGIMES[regime]
    o.zero_capacity = int(bool(zero_capacity))
    out = C.c_void_p()
    rc = L.mpdhg_gen_scenario_ex(base, scenario_seed, C.byref(o), C.byref(out))
    L.mpdhg_lp_free(base)
    _abi.check(rc, "gen_scenario_ex")
    return _owned_to_problem(out, f"{kind}_{'x'.join(map(str, size))}_s{base_seed}_{regime}_{scenario_seed}")


# ---- MPS / QPS files (io/mps_io.hpp:10-13) --------------------------------
def read_mps(filepath: str) -> Problem:
    """io::read_mps (mps_io.cpp:62-409), parsed by libmpdhg: the LP, name,
    sense, objective constant, row/column names, variable types and, for QPS
    files, Q.  Raises RuntimeError with the reference's message on a file that
    cannot be opened, a malformed number or contradictory bounds."""
    L = _abi.lib()
    ptr = C.c_void_p()
    _abi.check(L.mpdhg_read_mps(str(filepath).encode(), C.byref(ptr)), "read_mps")
    try:
        meta = L.mpdhg_lp_meta_view(ptr).contents
        n = L.mpdhg_lp_view(ptr).contents.n
        m = L.mpdhg_lp_view(ptr).contents.m
        name = meta.name.decode()
        sense = ObjectiveSense(meta.sense)
        offset = meta.obj_offset
        rows = [meta.row_names[i].decode() for i in range(m)] if meta.row_names else []
        cols = [meta.col_names[j].decode() for j in range(n)] if meta.col_names else []
        vt = (np.ctypeslib.as_array(meta.var_type, shape=(n,)).astype(np.int8, copy=True)
              if meta.var_type and n else np.zeros(n, np.int8))
        Q = None
        if meta.has_quadratic:
            qn = meta.q_nnz
            cp = np.ctypeslib.as_array(meta.q_col_ptr, shape=(n + 1,)).astype(np.int64, copy=True)
            ri = (np.ctypeslib.as_array(meta.q_row_ind, shape=(qn,)).astype(np.int64, copy=True)
                  if qn else np.zeros(0, np.int64))
            qv = (np.ctypeslib.as_array(meta.q_vals, shape=(qn,)).astype(np.float64, copy=True)
                  if qn else np.zeros(0))
            Q = SparseMatrix.from_csc(n, n, cp, ri, qv)
        p = _owned_to_problem(ptr, name)
        ptr = None  # freed by _owned_to_problem
    finally:
        if ptr is not None:
            L.mpdhg_lp_free(ptr)
    p.sense, p.obj_offset, p.row_names, p.col_names, p.var_types, p.Q = (
        sense, offset, rows, cols, vt, Q)
    return p


def _meta_view(problem: Problem):
    """mpdhg_lp_meta of a Problem, and the arrays it points into."""
    m, n = problem.num_rows(), problem.num_cols()
    meta = _abi.MpdhgLpMeta()
    meta.name = str(problem.name).encode()
    meta.sense = int(problem.sense)
    meta.obj_offset = float(problem.obj_offset)
    names = []
    if problem.row_names:
        if len(problem.row_names) < m:  # the reference pads with R<i>
            problem_rows = list(problem.row_names) + [f"R{i}" for i in range(len(problem.row_names), m)]
        else:
            problem_rows = list(problem.row_names)[:m]
        arr = (C.c_char_p * max(m, 1))(*[s.encode() for s in problem_rows])
        meta.row_names = C.cast(arr, C.POINTER(C.c_char_p))
        names.append(arr)
    if problem.col_names:
        if len(problem.col_names) < n:
            problem_cols = list(problem.col_names) + [f"C{j}" for j in range(len(problem.col_names), n)]
        else:
            problem_cols = list(problem.col_names)[:n]
        arr = (C.c_char_p * max(n, 1))(*[s.encode() for s in problem_cols])
        meta.col_names = C.cast(arr, C.POINTER(C.c_char_p))
        names.append(arr)
    vt = np.ascontiguousarray(problem.var_types if len(problem.var_types) == n else np.zeros(n), np.int8)
    meta.var_type = vt.ctypes.data_as(C.POINTER(C.c_int8))
    q = None
    if problem.Q is not None:
        Q = problem.Q
        q = (np.ascontiguousarray(Q.csc_col_ptr(), np.int64), np.ascontiguousarray(Q.csc_row_ind(), np.int64),
             np.ascontiguousarray(Q.csc_values(), np.float64))
        meta.has_quadratic = 1
        meta.q_nnz = Q.num_nonzeros()
        meta.q_col_ptr = q[0].ctypes.data_as(C.POINTER(C.c_int64))
        meta.q_row_ind = q[1].ctypes.data_as(C.POINTER(C.c_int64))
        meta.q_vals = q[2].ctypes.data_as(C.POINTER(C.c_double))
    return meta, (names, vt, q)


def write_mps(problem: Problem, filepath: str, free_format: bool = True) -> None:
    """io::write_mps (mps_io.cpp:411-578): the reference writer's text, byte
    for byte; free_format does not change the output, as there."""
    lp, keep = problem._c_view()
    meta, keep_meta = _meta_view(problem)
    _abi.check(_abi.lib().mpdhg_write_mps(str(filepath).encode(), C.byref(lp), C.byref(meta),
                                          int(bool(free_format))), "write_mps")


@dataclass
class CheckResult:
    """checker::CheckResult (checker.hpp:9-34)."""

    is_primal_feasible: bool = False
    max_row_violation: float = 0.0
    max_col_bound_violation: float = 0.0
    max_primal_residual: float = 0.0
    is_dual_feasible: bool = False
    max_dual_residual: float = 0.0
    is_complementary: bool = False
    max_complementarity_slack: float = 0.0
    is_integrality_satisfied: bool = False
    max_integrality_violation: float = 0.0
    evaluated_objective: float = 0.0
    objective_discrepancy: float = 0.0
    is_certificate_valid: bool = False
    certificate_violation: float = 0.0
    all_checks_passed: bool = False
    summary: str = ""


def check_solution(problem: Problem, solution: Solution, tol_primal: float = 1e-6,
                   tol_dual: float = 1e-6, tol_int: float = 1e-5, ray=None) -> CheckResult:
    """checker::check_solution (checker.cpp:9-268) in libmpdhg: independent FP64
    verification of a candidate solution.  The duals are read in the checker's
    sign (reduced cost = c - A^T y): pass a solution solved with dual_sign = 1,
    or negate the PDHG-sign row_duals first.  Empty vectors count as absent, as
    there.  ray: the certificate of an Infeasible / Unbounded status."""
    lp, keep = problem._c_view()
    meta, keep_meta = _meta_view(problem)
    dp = C.POINTER(C.c_double)
    inp = _abi.MpdhgCheckInput()
    inp.status = int(solution.status)
    arrs = []
    for name, v in (("x", solution.x), ("row_duals", solution.row_duals),
                    ("reduced_costs", solution.reduced_costs), ("ray", ray)):
        if v is not None and len(v):
            a = np.ascontiguousarray(v, np.float64)
            arrs.append(a)
            setattr(inp, name, a.ctypes.data_as(dp))
    want = {"x": problem.num_cols(), "row_duals": problem.num_rows(), "reduced_costs": problem.num_cols()}
    for name, k in want.items():  # a vector of the wrong length counts as absent, as there
        v = getattr(solution, name)
        if v is not None and len(v) and len(v) != k:
            setattr(inp, name, None)
    inp.primal_objective = float(solution.primal_objective)
    inp.ray_len = 0 if ray is None else len(ray)
    out = _abi.MpdhgCheckResult()
    _abi.check(_abi.lib().mpdhg_check_solution(C.byref(lp), C.byref(meta), C.byref(inp), tol_primal,
                                               tol_dual, tol_int, C.byref(out)), "check_solution")
    r = CheckResult()
    for name, _ in _abi.MpdhgCheckResult._fields_:
        v = getattr(out, name)
        if name == "summary":
            r.summary = v.decode(errors="replace")
        elif isinstance(getattr(r, name), bool):
            setattr(r, name, bool(v))
        else:
            setattr(r, name, float(v))
    return r


class MpsReader:
    """Static read/write pair, as py/refinery/parameterized_generator.py:191 calls it."""

    read = staticmethod(read_mps)

    @staticmethod
    def write(problem: Problem, filepath: str, free_format: bool = True) -> None:
        write_mps(problem, filepath, free_format)


__all__ = [
    "SparseMatrix", "Triplet", "Problem", "Solution", "SolutionStatus", "BasisStatus",
    "ObjectiveSense", "VariableType", "GpuSolver", "GpuPdhgConfig", "GpuMemoryInfo",
    "solve_gpu_pdhg", "solve_gpu_hybrid", "solve_batch", "identify_basis", "gen_random",
    "gen_staircase", "read_mps", "write_mps", "MpsReader", "polish", "PolishResult", "WarmStart",
    "check_solution", "CheckResult", "gen_scenario_ex", "REGIMES",
    "MPDHG_MODE_REFERENCE", "MPDHG_MODE_ENHANCED", "SIH_INFINITY",
]
