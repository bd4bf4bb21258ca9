"""The C ABI of libmpdhg.so without a GPU: exports, struct layouts, host logic.

* every function include/*.h declares is exported by the built library, and
  the ctypes binding declares exactly the functions of include/mpdhg.h;
* the ctypes structures have the size and field offsets the C compiler gives
  the header's structures (gcc on a generated probe);
* configuration defaults are the reference's (gpu_solver.hpp:12-22) and
  invalid extension values are rejected before any device work;
* without a device the solve entry points fail loudly (no CPU fallback);
* the seeded generators equal an independent Python restatement of their
  specification (SplitMix64-seeded xoshiro256**, Floyd sampling) bit for bit.
"""
import ctypes as C
import math
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import sih_solver as S
from sih_solver import _abi

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))
DECL = re.compile(r"^[A-Za-z_][\w\s\*]*?\b(mpdhg_\w+)\s*\(", re.MULTILINE)


def declared(path):
    text = re.sub(r"/\*.*?\*/", "", path.read_text(), flags=re.DOTALL)
    return sorted(set(DECL.findall(text)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(str(_abi.LIB_PATH))
    names = declared(ROOT / "include" / "mpdhg.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_abi.SIGNATURES) == names


def test_other_headers_are_exported_by_their_libraries():
    libs = {"mpdhg_vendor.h": ROOT / "refinery-optimizer_amd" / "lib" / "libmpdhg_vendor.so"}
    for h in HEADERS:
        if h.name == "mpdhg.h":
            continue
        assert h.name in libs, f"no library recorded for {h.name}"
        lib = C.CDLL(str(libs[h.name]))
        for n in declared(h):
            assert hasattr(lib, n), (h.name, n)
    from sih_solver import vendor

    assert sorted(vendor.SIGNATURES) == declared(ROOT / "include" / "mpdhg_vendor.h")


def _structs():
    from sih_solver import vendor

    return {"mpdhg_config": _abi.MpdhgConfig, "mpdhg_lp": _abi.MpdhgLp,
            "mpdhg_result": _abi.MpdhgResult, "mpdhg_adjoint_io": _abi.MpdhgAdjointIO,
            "mpdhg_lp_meta": _abi.MpdhgLpMeta, "mpdhg_polish_opts": _abi.MpdhgPolishOpts,
            "mpdhg_polish_result": _abi.MpdhgPolishResult,
            "mpdhg_vendor_result": vendor.MpdhgVendorResult,
            "mpdhg_vendor_stream_result": vendor.MpdhgVendorStreamResult,
            "mpdhg_vendor_gather_result": vendor.MpdhgVendorGatherResult}


def test_ctypes_layouts_match_the_header(tmp_path):
    STRUCTS = _structs()
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "mpdhg.h"', '#include "mpdhg_vendor.h"',
             "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n"):
        if line:
            a, b, v = line.split()
            got[(a, b)] = int(v)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)


def test_config_defaults_are_the_reference_defaults():
    c = _abi.MpdhgConfig()
    _abi.lib().mpdhg_config_default(C.byref(c))
    # GpuPdhgConfig{} (gpu_solver.hpp:12-22)
    assert c.max_iterations == 4000 and c.check_frequency == 25
    assert np.float32(c.tol_primal) == np.float32(1e-4) == np.float32(c.tol_dual) == np.float32(c.tol_gap)
    assert c.step_tau == 0.0 and c.step_sigma == 0.0
    assert c.enable_restart == 1 and c.enable_hybrid_polish == 1
    assert c.mode == _abi.MPDHG_MODE_REFERENCE and c.spmv_layout == 0 and c.spmv_chunk == 0
    py = S.GpuPdhgConfig()
    for name, _ in _abi.MpdhgConfig._fields_:
        assert math.isclose(float(getattr(py, name)), float(getattr(c, name)), rel_tol=1e-6), name


@pytest.mark.parametrize("field,value", [("spmv_chunk", 3000), ("spmv_layout", 2), ("mode", 7),
                                         ("check_frequency", 0), ("launch_graph", 2), ("spmv_block", 384),
                                         ("spmv_block", (128, 4096)), ("spmv_block", (512, 1024))])
def test_invalid_configuration_is_rejected(field, value):
    # mpdhg_solve with a NULL context validates the configuration before it
    # creates a temporary context, so this holds with or without a device.
    # A (block, chunk) value sets both: each is valid alone, the pair is no
    # compiled tile schedule.
    p = S.gen_random(20, 50, 4, 1)
    c = S.GpuPdhgConfig()
    if isinstance(value, tuple):
        c.spmv_block, c.spmv_chunk = value
    else:
        setattr(c, field, value)
    lp, keep = p._c_view()
    res = _abi.MpdhgResult()
    rc = _abi.lib().mpdhg_solve(None, C.byref(lp), C.byref(c._c()), C.byref(res))
    assert rc == -1  # MPDHG_E_ARG
    assert field in _abi.lib().mpdhg_last_error().decode() or "mode" in _abi.lib().mpdhg_last_error().decode()


def test_no_device_fails_loudly():
    if S.GpuSolver.is_cuda_available():
        pytest.skip("a HIP device is visible")
    p = S.gen_random(20, 50, 4, 1)
    with pytest.raises(_abi.MpdhgError):
        S.solve_gpu_pdhg(p)
    with pytest.raises(_abi.MpdhgError):
        S.solve_batch([p], S.GpuPdhgConfig.enhanced())
    with pytest.raises(_abi.MpdhgError):
        S.solve_gpu_hybrid(p)  # the PDHG stage needs the device; no CPU stand-in
    assert S.GpuSolver.get_vram_info().total_bytes == 0


def test_no_device_cpu_fallback_when_asked():
    # cpu_fallback = 1 reproduces GpuSolver::solve_pdhg without a device
    # (gpu_solver.cpp:66-71): the host FP64 simplex solves the LP.  toy01's
    # optimum is x* = (1.5, 2.5), objective -6.5 (toy_suite_metadata.json:5-12);
    # both rows are tight with simplex duals (-0.5, -0.5), i.e. (0.5, 0.5) in
    # the PDHG sign (reduced cost = c + A^T y = 0 on both basic columns).
    if S.GpuSolver.is_cuda_available():
        pytest.skip("a HIP device is visible")
    from lpfix import toy01

    p = toy01()
    for dual_sign, y_expect in ((0, [0.5, 0.5]), (1, [-0.5, -0.5])):
        s = S.solve_gpu_pdhg(p, S.GpuPdhgConfig(cpu_fallback=1, dual_sign=dual_sign))
        assert s.status == S.SolutionStatus.Optimal
        assert np.allclose(s.x, [1.5, 2.5], atol=1e-12) and math.isclose(s.primal_objective, -6.5)
        assert np.allclose(s.row_duals, y_expect, atol=1e-12)
        assert np.allclose(s.reduced_costs, 0.0, atol=1e-12) and np.allclose(s.slack, [4.0, 9.0])
        assert s.rel_kkt < 1e-12
    q = S.gen_random(30, 72, 4, 3)
    a, b = S.solve_batch([q, p], S.GpuPdhgConfig.enhanced(cpu_fallback=1))
    assert a.status == b.status == S.SolutionStatus.Optimal and a.rel_kkt < 1e-9
    h = S.solve_gpu_hybrid(p, S.GpuPdhgConfig(cpu_fallback=1))
    assert h.status == S.SolutionStatus.Optimal and math.isclose(h.primal_objective, -6.5)


# ---- generator restatement ---------------------------------------------------
M64 = (1 << 64) - 1


class Xoshiro:
    """xoshiro256** seeded through SplitMix64 (Blackman & Vigna)."""

    def __init__(self, seed):
        st = seed & M64
        self.s = []
        for _ in range(4):
            st = (st + 0x9E3779B97F4A7C15) & M64
            z = st
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
            self.s.append(z ^ (z >> 31))

    @staticmethod
    def rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & M64

    def next(self):
        s = self.s
        r = (self.rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self.rotl(s[3], 45)
        return r

    def u01(self):
        return (self.next() >> 11) * 2.0 ** -53

    def uni(self, a, b):
        return a + (b - a) * self.u01()

    def below(self, k):
        return self.next() % k

    def normal(self):
        u1 = self.u01()
        u2 = self.u01()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(6.283185307179586 * u2)

    def sample(self, n, k):  # Floyd: k distinct values of [0, n), ascending
        out = []
        for t in range(n - k, n):
            r = self.below(t + 1)
            out.append(t if r in out else r)
        return sorted(out)


def random_lp_spec(m, n, k, seed):
    """cfg0 generator: 8 (k) distinct rows per column, U[-1,1] values, 0 <= x <= 10,
    b = A x0 with x0 ~ U[0,10], c = A^T y0 + s, y0 ~ N(0,1), s ~ U[0,1] with
    floor(0.4 n) entries zeroed."""
    g = Xoshiro(seed)
    cp, ri, v = [0], [], []
    for _ in range(n):
        for r in g.sample(m, k):
            ri.append(r)
            v.append(g.uni(-1.0, 1.0))
        cp.append(len(ri))
    x0 = [g.uni(0.0, 10.0) for _ in range(n)]
    b = [0.0] * m
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            b[ri[p]] += v[p] * x0[j]
    y0 = [g.normal() for _ in range(m)]
    s = [g.u01() for _ in range(n)]
    for j in g.sample(n, (n * 4) // 10):
        s[j] = 0.0
    c = []
    for j in range(n):
        acc = 0.0
        for p in range(cp[j], cp[j + 1]):
            acc += v[p] * y0[ri[p]]
        c.append(acc + s[j])
    return np.array(cp), np.array(ri), np.array(v), np.array(b), np.array(c)


@pytest.mark.parametrize("m,n,k,seed", [(12, 30, 4, 0), (40, 96, 8, 7), (5, 9, 8, 3)])
def test_random_generator_matches_its_specification(m, n, k, seed):
    p = S.gen_random(m, n, k, seed)
    cp, ri, v, b, c = random_lp_spec(m, n, min(k, m), seed)
    A = p.A
    assert np.array_equal(A.csc_col_ptr(), cp) and np.array_equal(A.csc_row_ind(), ri)
    assert np.array_equal(A.csc_values(), v)
    assert np.array_equal(p.row_lower, b) and np.array_equal(p.row_upper, b)
    assert np.array_equal(p.c, c)
    assert np.all(p.col_lower == 0.0) and np.all(p.col_upper == 10.0)


def test_staircase_generator_structure():
    T, R, Cc = 3, 40, 96
    p = S.gen_staircase(T, R, Cc, 5)
    P, Q = R // 2, (3 * R) // 10
    NC = Cc - 2 * P
    assert (p.num_rows(), p.num_cols()) == (T * R, T * Cc)
    A = p.A
    cp, ri, v = A.csc_col_ptr(), A.csc_row_ind(), A.csc_values()
    for t in range(T):
        r0, c0 = t * R, t * Cc
        assert np.all(p.row_lower[r0:r0 + P] == 0) and np.all(p.row_upper[r0:r0 + P] == 0)
        assert np.all(np.isneginf(p.row_lower[r0 + P:r0 + R]))
        for j in range(c0, c0 + NC):  # crude: yields <= 1 in total, one unit row
            rows, vals = ri[cp[j]:cp[j + 1]], v[cp[j]:cp[j + 1]]
            y = vals[rows < r0 + P]
            assert 4 <= len(y) <= 8 and y.sum() <= 1.0 + 1e-12 and np.all(y > 0)
            assert np.sum(rows >= r0 + P + Q) == 1
            assert 40 <= p.c[j] <= 90
        for j in range(c0 + NC, c0 + NC + P):  # sales: revenue, -1 in the balance
            assert -120 <= p.c[j] <= -60 and v[cp[j]] == -1.0
        for j in range(c0 + NC + P, c0 + Cc):  # inventory: -1 now, +1 next block
            rows, vals = ri[cp[j]:cp[j + 1]], v[cp[j]:cp[j + 1]]
            assert vals[0] == -1.0 and rows[0] == r0 + (j - c0 - NC - P)
            assert (len(rows) == 2) == (t + 1 < T)
    assert np.all(p.col_lower == 0) and np.all((0 <= p.col_upper) & (p.col_upper <= 1000))


def test_scenario_generator_shocks_costs_and_zeroes_one_capacity():
    base = S.gen_staircase(2, 40, 96, 1)
    sc = S.gen_scenario("staircase", (2, 40, 96), 1, 9, 0.3, True)
    ratio = sc.c / base.c
    assert np.all(ratio > 0) and not np.allclose(ratio, 1.0)
    changed = np.flatnonzero(sc.row_upper != base.row_upper)
    assert len(changed) == 1 and sc.row_upper[changed[0]] == 0.0
    assert np.array_equal(sc.A.csc_values(), base.A.csc_values())


def test_crossover_and_mps_entry_points_reject_bad_arguments(tmp_path):
    L = _abi.lib()
    res = _abi.MpdhgPolishResult()
    assert L.mpdhg_polish(None, None, None, None, None, C.byref(res)) == -1  # NULL lp
    p = S.gen_random(20, 50, 4, 1)
    lp, keep = p._c_view()
    o = _abi.MpdhgPolishOpts()
    L.mpdhg_polish_opts_default(C.byref(o))
    assert (o.max_iterations, o.primal_tol, o.dual_tol, o.refactor_every) == (-1, 1e-6, 1e-6, 100)
    o.primal_tol = 0.0
    assert L.mpdhg_polish(C.byref(lp), None, None, None, C.byref(o), C.byref(res)) == -1
    o.primal_tol = 1e-6
    # col_ptr must span exactly nnz entries: a short count is refused before any read
    lp.nnz -= 1
    assert L.mpdhg_polish(C.byref(lp), None, None, None, C.byref(o), C.byref(res)) == -1
    assert "col_ptr must start at 0 and end at nnz" in L.mpdhg_last_error().decode()
    lp.nnz += 1
    assert L.mpdhg_polish(C.byref(lp), None, None, None, C.byref(o), C.byref(res)) == 0
    assert L.mpdhg_solve_hybrid(None, None, None, None, None, C.byref(res)) == -1
    assert L.mpdhg_read_mps(None, None) == -1
    out = C.c_void_p()
    assert L.mpdhg_read_mps(str(tmp_path / "none.mps").encode(), C.byref(out)) == -1
    assert "Could not open MPS file" in L.mpdhg_last_error().decode()
    assert L.mpdhg_write_mps(None, C.byref(lp), None, 1) == -1
    bad = tmp_path / "bad.mps"
    lp.nnz -= 1
    assert L.mpdhg_write_mps(str(bad).encode(), C.byref(lp), None, 1) == -1
    assert "col_ptr must start at 0 and end at nnz" in L.mpdhg_last_error().decode()
    lp.nnz += 1
    rows = np.ctypeslib.as_array(lp.row_ind, shape=(lp.nnz,))
    saved = int(rows[3])
    rows[3] = 20  # one past the last row
    assert L.mpdhg_write_mps(str(bad).encode(), C.byref(lp), None, 1) == -1
    assert "row index out of range" in L.mpdhg_last_error().decode()
    rows[3] = saved
    assert not bad.exists()
    # a NULL meta writes a nameless minimisation LP (rows R0.., columns C0..)
    path = tmp_path / "anon.mps"
    assert L.mpdhg_write_mps(str(path).encode(), C.byref(lp), None, 1) == 0
    q = S.read_mps(str(path))
    assert q.name == "problem" and q.row_names[0] == "R0" and q.col_names[0] == "C0"
    assert (q.num_rows(), q.num_cols()) == (20, 50)


def test_integration_shim_compiles_against_the_reference_headers(tmp_path):
    """The C++ shim INTEGRATION.md gives for GpuSolver (solve_pdhg, solve_hybrid,
    queries) compiles against the reference's own headers and include/mpdhg.h."""
    import re
    ref_inc = Path("/root/reference/core/include")
    if not (ref_inc / "sih" / "gpu" / "gpu_solver.hpp").exists():
        pytest.skip("reference headers not present")
    text = (ROOT / "INTEGRATION.md").read_text()
    shim = re.findall(r"```cpp\n(.*?)```", text, re.S)[0]
    src = tmp_path / "shim.cpp"
    src.write_text(shim)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", str(ref_inc), "-I", str(ROOT / "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_gather_probe_checks_its_arguments():
    # rejected before any device call, so this runs without a GPU
    code = f"""
import sys
sys.path.insert(0, {str(ROOT / "refinery-optimizer_amd")!r})
from sih_solver import vendor
for args in ((0.0, 8), (-1.0, 8), (65.0, 8), (1.0, 0), (1.0, 5000)):
    try:
        vendor.gather(args[0], per_xcd=args[1], reps=1)
    except RuntimeError as e:
        assert "code -1)" in str(e), str(e)
    else:
        raise SystemExit(f"accepted {{args}}")
print("ok")
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-2000:]
