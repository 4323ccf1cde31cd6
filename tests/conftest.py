import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "parallel-string-sorting_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: large inputs")


def load_ps5():
    spec = importlib.util.spec_from_file_location("ps5", os.path.join(PKG, "ps5.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ps5():
    return load_ps5()


@pytest.fixture(scope="session")
def oracle():
    import ps5_oracle
    ps5_oracle.load()
    return ps5_oracle


@pytest.fixture(scope="session")
def ctx(ps5):
    c = ps5.Context(0)
    yield c
    c.close()
