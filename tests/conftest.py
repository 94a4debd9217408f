import gc
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP engine through the C-ABI)"
    )
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP engine library, loaded through ctypes (fails loudly)."""
    from fluidsim_amd import _lib

    return _lib.get_lib()


@pytest.fixture(autouse=True)
def _release_device_memory():
    """After each test: collect the simulations left in reference cycles
    (their plans hold device workspaces outside torch's allocator) and hand
    torch's cached blocks back, so a later test can allocate a large plan."""
    yield
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
