import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu_ctx():
    """Context on cuda:0; fails (does not skip) when the HIP library is missing."""
    import torch

    import clear_amd as ca

    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    ca.lib()  # raises if libclear_amd.so was not built
    ctx = ca.Context(0)
    yield ctx
    ctx.close()
