import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "softx-d-21-00092_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.LIB):
        oracle.build(ref=False)
    return oracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle
    if not os.path.exists(oracle.REF_LIB):
        if os.path.isdir("/root/reference"):
            oracle.build(ref=True)
        else:
            pytest.skip("reference build (oracle/_ref) not available")
    return oracle.Reference()


@pytest.fixture(scope="session")
def jw():
    """The HIP library's Python binding; GPU tests only."""
    import jimwlk
    return jimwlk
