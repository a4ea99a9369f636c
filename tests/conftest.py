import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

__graft_entry__.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdonofhip.so)")
    config.addinivalue_line("markers", "slow: larger CPU oracle cases")


@pytest.fixture(scope="session")
def oracle():
    return __graft_entry__.load_oracle()


@pytest.fixture(scope="session")
def dn():
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def gpu_ctx():
    from donof_amd import api
    return api.default_context()
