import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO / "mossco-code_amd", REPO / "oracle", REPO / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsedcol.so on cuda:0)")
    config.addinivalue_line("markers", "slow: longer CPU runs")


@pytest.fixture(scope="session")
def gpu_ready():
    """Initialise the GPU once; GPU tests fail (never fall back) without it."""
    import torch

    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    torch.cuda.init()
    from sedcol import capi

    capi.lib()
    return True
