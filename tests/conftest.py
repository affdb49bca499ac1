import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle():
    from tests import _oracle
    return _oracle.load()


@pytest.fixture(scope="session")
def gpu_ctx():
    """One device context shared by the GPU tests of a session."""
    from dynasor_plus_amd import GpuContext
    ctx = GpuContext(0)
    yield ctx
    ctx.close()


def golden(name):
    import numpy as np
    return np.load(GOLDEN / name, allow_pickle=False)
