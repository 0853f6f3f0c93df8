import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def gpu():
    """Initialises the library on device 0 (GPU tests only)."""
    import torch
    from hipacc_amd import operators

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    operators.init(0)
    return torch.device("cuda:0")
