import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    # One HIP runtime per process: PyTorch's must be loaded before
    # libcabanadem_hip.so initialises the device, or torch sees no GPU later
    # in the session (the surrogate tests move data through torch tensors).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; GPU tests fail (not skip) when it cannot be loaded."""
    from cabanadem_amd import abi

    lib = abi.lib()
    if abi.device_count() < 1:
        pytest.fail("no HIP device visible to libcabanadem_hip.so")
    return lib
