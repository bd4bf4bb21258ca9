"""bench.py's roofline bookkeeping (CPU only): the kernel label follows the
stream-size rule of the engine, every label the sweep can report has a
committed traffic figure, and those figures are the PMC sums the traffic
summary states."""
import importlib.util
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_kernel_label_follows_the_stream_size_rule():
    b = _bench()
    limit = 256 * 2**20 // 16  # nonzeros at the 256 MiB threshold (16 B per nonzero)
    assert b.enh_kernels(limit) == ("k_enh_primal", "k_enh_dual")
    assert b.enh_kernels(limit + 1) == ("k_enh_primal_nt", "k_enh_dual_nt")
    assert b.enh_kernels(40_639_151) == ("k_enh_primal_nt", "k_enh_dual_nt")  # sweep512


def test_engine_and_bench_use_the_same_threshold():
    src = (ROOT / "refinery-optimizer_amd" / "csrc" / "engine_enh.cpp").read_text()
    m = re.search(r"constexpr double kNtStreamBytes = ([0-9.]+) \* 1024 \* 1024;", src)
    assert m and float(m.group(1)) * 2**20 == _bench().NT_STREAM_BYTES


def test_committed_traffic_matches_the_counter_table():
    b = _bench()
    text = b.TRAFFIC_FILE.read_text()
    for k in ("k_enh_primal_nt", "k_enh_dual_nt", "k_enh_primal", "k_enh_dual"):
        assert b.measured_traffic(k) is not None, k
    # 2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE, KiB -> bytes, from the table rows
    for k in ("k_enh_primal_nt", "k_enh_dual_nt"):
        row = re.search(rf"^\| `{k}<[^|]*\| ([0-9,.]+) \| ([0-9,.]+) \|", text, re.MULTILINE)
        fetch, write = (float(v.replace(",", "")) for v in row.groups())
        assert abs(b.measured_traffic(k) - (2.0 * fetch + write) * 1024) < 1024, k
    assert b.measured_traffic("k_no_such_kernel") is None


def test_comparison_samples_draw_both_families_alike():
    # The sweep alternates its generator families, so the plain, CPU and
    # vertex samples are stratified: every k-th LP of each family.
    import sih_solver  # noqa: F401  (conftest puts the package on the path)
    from sih_solver import workloads as W

    b = _bench()
    fam = [sp[0] for sp in W.sweep_specs(512)]
    for step, per in ((16, 32), (8, 64)):
        idx = b.strata(512, step)
        assert len(idx) == 2 * per and len(set(idx)) == len(idx)
        assert sum(fam[q] == "random" for q in idx) == sum(fam[q] == "staircase" for q in idx) == per


class _Lp:
    def __init__(self, m, n, nnz):
        self.m, self.n, self.nnz = m, n, nnz

    def num_rows(self):
        return self.m

    def num_cols(self):
        return self.n

    def num_nonzeros(self):
        return self.nnz


def test_fused_bytes_follow_the_iteration_formula():
    b = _bench()
    probs = [_Lp(1000, 2400, 19_000), _Lp(200, 480, 3_100), _Lp(0, 5, 0)]
    nnz, n, m, k = 22_100, 2885, 1200, 3
    sp, sd = b.fused_bytes(probs, False)
    vp, vd = b.fused_bytes(probs, True)
    # 16 nnz + 32 n + 24 m per iteration (scalar steps), + 4 n + 4 m with
    # per-column / per-row steps; the +4 per LP is each pointer array's end
    assert sp + sd == 16 * nnz + 32 * n + 24 * m + 4 * 2 * k
    assert vp + vd == sp + sd + 4 * n + 4 * m
    assert (vp - sp, vd - sd) == (4 * n, 4 * m)
    # the engine's reference-mode count per LP (engine.cpp) is the vector-step split
    for q in probs:
        ep = 8.0 * q.nnz + 4.0 * (q.n + 1) + 4.0 * q.m + 4.0 * 7 * q.n
        ed = 8.0 * q.nnz + 4.0 * (q.m + 1) + 4.0 * q.n + 4.0 * 5 * q.m
        assert b.fused_bytes([q], True) == (ep, ed)


def test_roofline_reports_formula_and_engine_bytes():
    b = _bench()

    class Sol:
        kernel_launches_csc = 10
        kernel_ms_spmv_csc = 1.0   # 100 us per launch
        kernel_ms_spmv_csr = 0.5   # 50 us
        bytes_csc_launch = 5.2e8
        bytes_csr_launch = 4.2e8
        csr_block, csr_chunk, csc_block, csc_chunk, tune_ms = 512, 8192, 512, 4096, 0.0

    rf, spmv = b.roofline(Sol(), ("kp", "kd"), traffic=False, algo=(5.0e8, 4.0e8))
    assert rf["kernel"] == "kp" and rf["us_per_launch"] == 100.0
    assert abs(rf["achieved"] - 5000.0) < 1e-9 and abs(rf["frac"] - 0.625) < 1e-12
    assert abs(rf["engine_achieved"] - 5200.0) < 1e-9
    assert rf["traffic"] is None
    assert abs(spmv["csr_fused_gbs"] - 8000.0) < 1e-9
