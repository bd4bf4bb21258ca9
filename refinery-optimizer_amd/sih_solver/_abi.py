"""ctypes binding of libmpdhg.so (include/mpdhg.h).

The shared library is the product; this module only declares its C types.
Loading fails loudly when the library is missing: there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("MPDHG_LIB", _HERE.parent / "lib" / "libmpdhg.so"))

MPDHG_MODE_REFERENCE = 0
MPDHG_MODE_ENHANCED = 1


class MpdhgConfig(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32),
        ("tol_primal", C.c_float),
        ("tol_dual", C.c_float),
        ("tol_gap", C.c_float),
        ("check_frequency", C.c_int32),
        ("step_tau", C.c_float),
        ("step_sigma", C.c_float),
        ("enable_restart", C.c_int32),
        ("enable_hybrid_polish", C.c_int32),
        ("mode", C.c_int32),
        ("tol_kkt", C.c_double),
        ("max_refinements", C.c_int32),
        ("reflection", C.c_float),
        ("ruiz_iterations", C.c_int32),
        ("profile", C.c_int32),
        ("dual_sign", C.c_int32),
        ("refine_ratio", C.c_double),
        ("spmv_layout", C.c_int32),
        ("spmv_chunk", C.c_int32),
        ("launch_graph", C.c_int32),
        ("spmv_block", C.c_int32),
        ("spmv_tune", C.c_int32),
        ("cpu_fallback", C.c_int32),
        ("vertex_check", C.c_int32),
        ("spmv_bands", C.c_int32),
    ]


class MpdhgLp(C.Structure):
    _fields_ = [
        ("m", C.c_int64),
        ("n", C.c_int64),
        ("nnz", C.c_int64),
        ("col_ptr", C.POINTER(C.c_int64)),
        ("row_ind", C.POINTER(C.c_int64)),
        ("vals", C.POINTER(C.c_double)),
        ("c", C.POINTER(C.c_double)),
        ("col_lower", C.POINTER(C.c_double)),
        ("col_upper", C.POINTER(C.c_double)),
        ("row_lower", C.POINTER(C.c_double)),
        ("row_upper", C.POINTER(C.c_double)),
    ]


class MpdhgResult(C.Structure):
    _fields_ = [
        ("x", C.POINTER(C.c_double)),
        ("y", C.POINTER(C.c_double)),
        ("reduced_costs", C.POINTER(C.c_double)),
        ("row_activity", C.POINTER(C.c_double)),
        ("status", C.c_int32),
        ("restarts", C.c_int32),
        ("refinements", C.c_int32),
        ("checks", C.c_int32),
        ("iterations", C.c_int64),
        ("primal_objective", C.c_double),
        ("dual_objective", C.c_double),
        ("rel_primal", C.c_double),
        ("rel_dual", C.c_double),
        ("rel_gap", C.c_double),
        ("rel_kkt", C.c_double),
        ("ref_primal", C.c_double),
        ("ref_dual", C.c_double),
        ("time_wall_sec", C.c_double),
        ("solve_sec", C.c_double),
        ("kernel_ms_spmv_csr", C.c_double),
        ("kernel_ms_spmv_csc", C.c_double),
        ("kernel_launches_csr", C.c_int64),
        ("kernel_launches_csc", C.c_int64),
        ("bytes_csr_launch", C.c_double),
        ("bytes_csc_launch", C.c_double),
        ("step_size", C.c_double),
        ("primal_weight", C.c_double),
        ("csr_block", C.c_int32),
        ("csr_chunk", C.c_int32),
        ("csc_block", C.c_int32),
        ("csc_chunk", C.c_int32),
        ("tune_ms", C.c_double),
        ("csr_bands", C.c_int32),
        ("vertex_status", C.c_int32),
        ("vertex_objective", C.c_double),
        ("vertex_rel_error", C.c_double),
    ]


class MpdhgWarmStart(C.Structure):
    _fields_ = [
        ("x", C.POINTER(C.c_double)),
        ("y", C.POINTER(C.c_double)),
        ("primal_weight", C.c_double),
        ("step_size", C.c_double),
    ]


class MpdhgScenarioOpts(C.Structure):
    _fields_ = [
        ("shock_sigma", C.c_double),
        ("price_std", C.c_double),
        ("supply_std", C.c_double),
        ("capacity_std", C.c_double),
        ("demand_std", C.c_double),
        ("regime", C.c_int32),
        ("zero_capacity", C.c_int32),
    ]


class MpdhgCheckInput(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("x", C.POINTER(C.c_double)),
        ("row_duals", C.POINTER(C.c_double)),
        ("reduced_costs", C.POINTER(C.c_double)),
        ("primal_objective", C.c_double),
        ("ray", C.POINTER(C.c_double)),
        ("ray_len", C.c_int64),
    ]


class MpdhgCheckResult(C.Structure):
    _fields_ = [
        ("is_primal_feasible", C.c_int32),
        ("max_row_violation", C.c_double),
        ("max_col_bound_violation", C.c_double),
        ("max_primal_residual", C.c_double),
        ("is_dual_feasible", C.c_int32),
        ("max_dual_residual", C.c_double),
        ("is_complementary", C.c_int32),
        ("max_complementarity_slack", C.c_double),
        ("is_integrality_satisfied", C.c_int32),
        ("max_integrality_violation", C.c_double),
        ("evaluated_objective", C.c_double),
        ("objective_discrepancy", C.c_double),
        ("is_certificate_valid", C.c_int32),
        ("certificate_violation", C.c_double),
        ("all_checks_passed", C.c_int32),
        ("summary", C.c_char * 256),
    ]


class MpdhgAdjointIO(C.Structure):
    _fields_ = [
        ("x", C.POINTER(C.c_double)),
        ("y", C.POINTER(C.c_double)),
        ("w", C.POINTER(C.c_double)),
        ("dl_dc", C.POINTER(C.c_double)),
        ("dl_db", C.POINTER(C.c_double)),
        ("dl_dl", C.POINTER(C.c_double)),
        ("dl_du", C.POINTER(C.c_double)),
        ("basic_cols", C.c_int32),
        ("active_rows", C.c_int32),
        ("iterations", C.c_int32),
        ("residual", C.c_double),
    ]


class MpdhgLpMeta(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("sense", C.c_int32),
        ("obj_offset", C.c_double),
        ("row_names", C.POINTER(C.c_char_p)),
        ("col_names", C.POINTER(C.c_char_p)),
        ("var_type", C.POINTER(C.c_int8)),
        ("has_quadratic", C.c_int32),
        ("q_nnz", C.c_int64),
        ("q_col_ptr", C.POINTER(C.c_int64)),
        ("q_row_ind", C.POINTER(C.c_int64)),
        ("q_vals", C.POINTER(C.c_double)),
    ]


class MpdhgPolishOpts(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int64),
        ("primal_tol", C.c_double),
        ("dual_tol", C.c_double),
        ("obj_offset", C.c_double),
        ("refactor_every", C.c_int32),
        ("time_limit_sec", C.c_double),
    ]


class MpdhgPolishResult(C.Structure):
    _fields_ = [
        ("x", C.POINTER(C.c_double)),
        ("y", C.POINTER(C.c_double)),
        ("reduced_costs", C.POINTER(C.c_double)),
        ("row_activity", C.POINTER(C.c_double)),
        ("col_basis_out", C.POINTER(C.c_int8)),
        ("row_basis_out", C.POINTER(C.c_int8)),
        ("status", C.c_int32),
        ("warm_start", C.c_int32),
        ("iterations", C.c_int64),
        ("refactorizations", C.c_int64),
        ("repaired", C.c_int64),
        ("bound_flips", C.c_int64),
        ("degenerate_pivots", C.c_int64),
        ("push_steps", C.c_int64),
        ("objective", C.c_double),
        ("primal_infeasibility", C.c_double),
        ("time_sec", C.c_double),
    ]


# name -> (restype, argtypes); exactly the symbols include/mpdhg.h declares.
_VP = C.c_void_p
SIGNATURES = {
    "mpdhg_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mpdhg_mem_info": (C.c_int, [C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                 C.POINTER(C.c_size_t)]),
    "mpdhg_last_error": (C.c_char_p, []),
    "mpdhg_config_default": (None, [C.POINTER(MpdhgConfig)]),
    "mpdhg_create": (C.c_int, [C.c_int, C.POINTER(_VP)]),
    "mpdhg_destroy": (None, [_VP]),
    "mpdhg_solve": (C.c_int, [_VP, C.POINTER(MpdhgLp), C.POINTER(MpdhgConfig),
                              C.POINTER(MpdhgResult)]),
    "mpdhg_solve_batch": (C.c_int, [_VP, C.POINTER(MpdhgLp), C.c_int, C.POINTER(MpdhgConfig),
                                    C.POINTER(MpdhgResult)]),
    "mpdhg_solve_noalloc": (C.c_int, [_VP, C.POINTER(MpdhgLp), C.POINTER(MpdhgConfig),
                                      C.POINTER(MpdhgResult)]),
    "mpdhg_solve_warm": (C.c_int, [_VP, C.POINTER(MpdhgLp), C.c_int, C.POINTER(MpdhgConfig),
                                   C.POINTER(MpdhgWarmStart), C.POINTER(MpdhgResult)]),
    "mpdhg_adjoint_batch": (C.c_int, [_VP, C.POINTER(MpdhgLp), C.c_int, C.POINTER(MpdhgAdjointIO),
                                      C.c_double, C.c_double, C.c_int]),
    "mpdhg_identify_basis": (C.c_int, [C.POINTER(MpdhgLp), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int8), C.POINTER(C.c_int8)]),
    "mpdhg_gen_random": (C.c_int, [C.c_int64, C.c_int64, C.c_int32, C.c_uint64, C.POINTER(_VP)]),
    "mpdhg_gen_staircase": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                      C.POINTER(_VP)]),
    "mpdhg_gen_scenario": (C.c_int, [_VP, C.c_uint64, C.c_double, C.c_int32, C.POINTER(_VP)]),
    "mpdhg_scenario_opts_default": (None, [C.POINTER(MpdhgScenarioOpts)]),
    "mpdhg_gen_scenario_ex": (C.c_int, [_VP, C.c_uint64, C.POINTER(MpdhgScenarioOpts), C.POINTER(_VP)]),
    "mpdhg_lp_view": (C.POINTER(MpdhgLp), [_VP]),
    "mpdhg_lp_row_kind": (C.POINTER(C.c_int8), [_VP]),
    "mpdhg_lp_free": (None, [_VP]),
    "mpdhg_polish_opts_default": (None, [C.POINTER(MpdhgPolishOpts)]),
    "mpdhg_polish": (C.c_int, [C.POINTER(MpdhgLp), C.POINTER(C.c_int8), C.POINTER(C.c_int8),
                               C.POINTER(C.c_double), C.POINTER(MpdhgPolishOpts),
                               C.POINTER(MpdhgPolishResult)]),
    "mpdhg_solve_hybrid": (C.c_int, [_VP, C.POINTER(MpdhgLp), C.POINTER(MpdhgConfig),
                                     C.POINTER(MpdhgPolishOpts), C.POINTER(MpdhgResult),
                                     C.POINTER(MpdhgPolishResult)]),
    "mpdhg_read_mps": (C.c_int, [C.c_char_p, C.POINTER(_VP)]),
    "mpdhg_lp_meta_view": (C.POINTER(MpdhgLpMeta), [_VP]),
    "mpdhg_write_mps": (C.c_int, [C.c_char_p, C.POINTER(MpdhgLp), C.POINTER(MpdhgLpMeta), C.c_int32]),
    "mpdhg_check_solution": (C.c_int, [C.POINTER(MpdhgLp), C.POINTER(MpdhgLpMeta), C.POINTER(MpdhgCheckInput),
                                       C.c_double, C.c_double, C.c_double, C.POINTER(MpdhgCheckResult)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libmpdhg.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"libmpdhg.so not found at {LIB_PATH}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (or make -C "
                "refinery-optimizer_amd/csrc)")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class MpdhgError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = lib().mpdhg_last_error().decode(errors="replace")
        super().__init__(f"{where} failed (code {code}): {msg}")
        self.code = code


def check(code: int, where: str) -> None:
    if code != 0:
        raise MpdhgError(code, where)
