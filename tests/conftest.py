import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "lb-cube_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def gpu():
    """The torch CUDA device; GPU tests fail loudly (no skip) when it is missing."""
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    import lbcube

    lbcube.lib()  # raises if the kernel library is not built
    return torch.device("cuda", 0)
