import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu_ctx():
    from raptor_amd import rt

    return rt.Context(0)


@pytest.fixture(scope="session")
def rt():
    from raptor_amd import rt as _rt

    return _rt


@pytest.fixture(scope="session")
def O():
    import oracle.oracle as _o

    return _o
