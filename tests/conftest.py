import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", os.path.join("cpm-ibc_amd", "python")):
    p = os.path.join(REPO, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through its C ABI)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


@pytest.fixture(scope="session")
def gpu_ctx():
    import cpmgpu
    ctx = cpmgpu.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture
def parallel_oracle():
    """Host threads for the long oracle runs (results do not depend on them)."""
    import orc
    orc.set_parallel(True, True, min(8, os.cpu_count() or 1))
    yield
    orc.set_parallel(False, False, 0)
