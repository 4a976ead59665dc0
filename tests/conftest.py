import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP renderer)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import py_oracle
    return py_oracle.lib()


@pytest.fixture(scope="session")
def renderer():
    """One HIP context for all GPU tests.  No skip: on a GPU box a missing
    library or device must fail loudly."""
    import clovers_amd
    r = clovers_amd.Renderer(0)
    yield r
    r.close()
