import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    import _oracle
    if not all(os.path.exists(f) for f in (_oracle.ORACLE_SO, _oracle.ORACLE_ULP_SO, _oracle.ORACLE_CR_SO)):
        _oracle.build_oracle()
    return _oracle
