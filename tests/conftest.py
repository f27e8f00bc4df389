import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); parity tests of the gfx950 casters")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch

    from robartsvtk_amd import abi

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    torch.cuda.init()
    ctx = abi.Context(0)
    yield ctx
    torch.cuda.synchronize()
    ctx.close()


@pytest.fixture(scope="session")
def nthreads():
    return max(1, min(16, os.cpu_count() or 1))
