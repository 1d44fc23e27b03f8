import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "hashpoint-code_amd"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def dirty_device_memory(request):
    """Fills free memory of torch's caching allocator with a junk byte
    pattern (0xA5) before each GPU test.  Outputs are allocated uninitialised,
    so an element a kernel fails to write then shows up as junk instead of
    keeping a value left by an earlier, identical call."""
    if request.node.get_closest_marker("gpu") is None:
        yield
        return
    import torch
    if not torch.cuda.is_available():
        yield
        return
    big = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0").fill_(0xA5)
    small = [torch.empty(256 << 10, dtype=torch.uint8, device="cuda:0").fill_(0xA5) for _ in range(32)]
    del big, small
    yield
