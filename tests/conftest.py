import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP allocator through libouro.so)")


@pytest.fixture(scope="session")
def gpu_lib():
    """libouro.so loaded on a machine with a GPU; a missing library is an error, never a skip."""
    import ouroboros_amd

    lib = ouroboros_amd.lib()
    return lib
