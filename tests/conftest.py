import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips GPU tests cleanly off-GPU; on a GPU box the native library must load."""
    if not _gpu_available():
        pytest.skip("no GPU")
    import aerof_dev_amd
    aerof_dev_amd.lib()  # raises if libaerofx.so is missing: no silent fallback
    import torch
    return torch.device("cuda:0")
