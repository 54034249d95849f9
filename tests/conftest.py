import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "h2factory.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through libh2g.so)")


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.fail("GPU test collected without a GPU (run with -m 'not gpu' on the CPU)")
    return 0


@pytest.fixture(autouse=True)
def _gpu_runtime_first(request):
    """GPU tests bring up torch's HIP runtime before the library's first call (torch does not
    come up in a process where libh2g already initialised the device)."""
    if request.node.get_closest_marker("gpu") is not None:
        request.getfixturevalue("gpu")
