import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "refinery-optimizer_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def gpu_available() -> bool:
    import sih_solver

    return sih_solver.GpuSolver.is_cuda_available()


@pytest.fixture(scope="session")
def need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device visible")
