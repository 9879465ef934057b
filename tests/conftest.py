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


@pytest.hookimpl(trylast=True)
def pytest_collection_modifyitems(config, items):
    """torch has to initialise the HIP runtime before libo3d is loaded into
    the process (loaded first, libo3d leaves torch seeing no device), so a
    run that selects GPU tests initialises it once, before any test runs,
    whatever the selection and order."""
    if any(item.get_closest_marker("gpu") for item in items):
        import torch
        torch.cuda.is_available()
