import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "dft-zig_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available() -> bool:
    try:
        import torch

        return bool(torch.cuda.is_available())
    except Exception:
        return False


@pytest.fixture(scope="session")
def dz():
    """The product binding; builds libdzhip.so in-tree when it is missing (CPU-side build)."""
    import dzhip
    from dzhip import build as dzbuild

    if not dzhip.LIB_PATH.exists():
        dzbuild.build()
    dzhip.L()
    return dzhip


@pytest.fixture(scope="session")
def ctx(dz):
    if not gpu_available():
        pytest.skip("no GPU")
    return dz.Context(0)
