import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def oracle_lib():
    import engines
    if not engines.oracle_available():
        pytest.fail("oracle/libne_oracle.so is not built (run __graft_entry__.build())")
    return engines


@pytest.fixture(scope="session")
def gpu_available():
    from nativeengine_amd import gpu
    gpu.load_library()   # raises loudly when the HIP library is missing
    return True
