import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def orc():
    from oracle import load_oracle
    return load_oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle import load_reference
    r = load_reference()
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree unavailable at build time)")
    return r
