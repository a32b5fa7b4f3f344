import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def orc():
    import oracle_lib

    return oracle_lib.oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle_lib

    r = oracle_lib.reference()
    if r is None:
        pytest.skip("reference harness (oracle/_ref) not built here")
    return r
