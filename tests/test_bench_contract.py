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
