"""ctypes binding of the bench-only rocSPARSE baseline (include/mpdhg_vendor.h).

bench/cusparse_pdhg_baseline.cu:60-255 counterpart: a fixed number of
reference PDHG iterations with rocSPARSE products, for comparison with the
engine's fused kernels, and ``stream()``, the device's streaming ceilings
(the practical HBM bandwidth of plain float4 read and copy kernels).  Not
used by any solve path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import Problem
from ._abi import MpdhgLp

LIB_PATH = Path(os.environ.get("MPDHG_VENDOR_LIB",
                               Path(__file__).resolve().parent.parent / "lib" / "libmpdhg_vendor.so"))
TRANSPOSE_OP = 0   # A^T y as the transposed product of the CSR (the reference's way)
EXPLICIT_CSC = 1   # A^T y with an explicit CSR of A^T


class MpdhgVendorResult(C.Structure):
    _fields_ = [("us_per_iter", C.c_double), ("us_spmv_t", C.c_double), ("us_spmv", C.c_double),
                ("objective", C.c_double)]


class MpdhgVendorStreamResult(C.Structure):
    _fields_ = [("read_gbs", C.c_double), ("read_nt_gbs", C.c_double), ("copy_gbs", C.c_double)]


class MpdhgVendorGatherResult(C.Structure):
    _fields_ = [("entries", C.c_int64), ("stream_us", C.c_double), ("gather_us", C.c_double),
                ("staged_us", C.c_double), ("stream_gbs", C.c_double)]


SIGNATURES = {
    "mpdhg_vendor_stream": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(MpdhgVendorStreamResult)]),
    "mpdhg_vendor_gather": (C.c_int, [C.c_int, C.c_double, C.c_int, C.POINTER(MpdhgVendorGatherResult)]),
    "mpdhg_vendor_pdhg": (C.c_int, [C.POINTER(MpdhgLp), C.c_int, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(MpdhgVendorResult)]),
    "mpdhg_vendor_last_error": (C.c_char_p, []),
}

_lib = None


def lib() -> C.CDLL:
    """Load libmpdhg_vendor.so.  Refuses a process that has imported torch:
    torch bundles its own HIP runtime and rocSPARSE, the dynamic linker then
    binds this library's rocSPARSE to torch's copy, and rocSPARSE fails on
    the other runtime's device state."""
    global _lib
    if _lib is None:
        import sys

        if "torch" in sys.modules:
            raise RuntimeError("the rocSPARSE baseline must run in a process that has not "
                               "imported torch (torch bundles its own ROCm libraries)")
        if not LIB_PATH.exists():
            raise RuntimeError(f"libmpdhg_vendor.so not found at {LIB_PATH}; build with "
                               "make -C refinery-optimizer_amd/csrc")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            getattr(h, name).restype = res
            getattr(h, name).argtypes = args
        _lib = h
    return _lib


@dataclass
class VendorRun:
    us_per_iter: float
    us_spmv_t: float
    us_spmv: float
    objective: float
    x: np.ndarray


def run(problem: Problem, iterations: int, mode: int = TRANSPOSE_OP) -> VendorRun:
    lp, keep = problem._c_view()
    x = np.zeros(problem.num_cols())
    r = MpdhgVendorResult()
    rc = lib().mpdhg_vendor_pdhg(C.byref(lp), iterations, mode,
                                 x.ctypes.data_as(C.POINTER(C.c_double)), C.byref(r))
    if rc != 0:
        raise RuntimeError(f"mpdhg_vendor_pdhg failed (code {rc}): "
                           f"{lib().mpdhg_vendor_last_error().decode(errors='replace')}")
    del keep
    return VendorRun(r.us_per_iter, r.us_spmv_t, r.us_spmv, r.objective, x)


def gather(stride: float, per_xcd: int = 620, reps: int = 20) -> dict:
    """Gather cost in the fused tile kernels' memory pattern
    (mpdhg_vendor_gather): mean µs per launch of the index/value stream of
    8 x per_xcd chunks of 8192 nonzeros alone, with sorted gathers `stride`
    vector entries apart, and with the LDS pair stage and row walk."""
    r = MpdhgVendorGatherResult()
    rc = lib().mpdhg_vendor_gather(per_xcd, stride, reps, C.byref(r))
    if rc != 0:
        raise RuntimeError(f"mpdhg_vendor_gather failed (code {rc}): "
                           f"{lib().mpdhg_vendor_last_error().decode(errors='replace')}")
    return {"entries": r.entries, "stream_us": r.stream_us, "gather_us": r.gather_us,
            "staged_us": r.staged_us, "stream_gbs": r.stream_gbs}


def stream(bytes_: int = 1 << 30, reps: int = 10) -> dict:
    """Streaming ceilings of device 0 (mpdhg_vendor_stream): best GB/s of a
    cached float4 read, a non-temporal read and a copy (read + write bytes)."""
    r = MpdhgVendorStreamResult()
    rc = lib().mpdhg_vendor_stream(bytes_, reps, C.byref(r))
    if rc != 0:
        raise RuntimeError(f"mpdhg_vendor_stream failed (code {rc}): "
                           f"{lib().mpdhg_vendor_last_error().decode(errors='replace')}")
    return {"read_gbs": r.read_gbs, "read_nt_gbs": r.read_nt_gbs, "copy_gbs": r.copy_gbs}
