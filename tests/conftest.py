import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # load what is built; build only when nothing is
    from fgrain_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from fgrain_amd import build as B
        B.build()
    from oracle import oracle as O
    if not os.path.exists(O._LIB_PATH):
        O.build()


@pytest.fixture(scope="session")
def engine():
    import fgrain_amd as F
    return F.GpuEngine.new()  # raises GpuError when no HIP device: GPU tests fail loudly
