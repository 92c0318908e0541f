import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def nm():
    """The product package (HIP library bindings)."""
    import __graft_entry__
    mod = __graft_entry__.package()
    mod.load()
    return mod


@pytest.fixture(scope="session")
def oracle_f64():
    from oracle import Oracle
    return Oracle("f64")


@pytest.fixture(scope="session")
def oracle_f32():
    from oracle import Oracle
    return Oracle("f32")
