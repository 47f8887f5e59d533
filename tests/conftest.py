import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on the GPU host)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def lib():
    import foundation_amd
    return foundation_amd.load()


@pytest.fixture(scope="session")
def oracle():
    from oracle import pt_oracle
    pt_oracle.load()
    return pt_oracle


@pytest.fixture(scope="session")
def ctx():
    """One device context for every GPU test (one process, one stream)."""
    from foundation_amd import pt
    c = pt.Context(0)
    yield c
    c.close()
