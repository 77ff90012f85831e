"""Shared test setup: import paths, the `gpu` marker, and GPU availability."""
import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
PKG = REPO / "ferebus-v7_amd"
for p in (str(REPO), str(PKG)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def _gpu_present() -> bool:
    try:
        import torch  # noqa: F401  (device discovery only)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    """Skips GPU tests where no GPU is visible; fails loudly if the library is missing."""
    if not _gpu_present():
        pytest.skip("no GPU visible")
    import ferebus_amd
    ferebus_amd.load()      # raises if libferebus_hip.so was not built
    return ferebus_amd
