import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def tm():
    """The HIP library through its C-ABI (fails loudly when it is not built)."""
    import timothy_amd

    return timothy_amd
