import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(autouse=True)
def _reset_oracle_mutation():
    """No test may leak a planted oracle mistake into the next one."""
    yield
    if "oracle" in sys.modules:
        try:
            sys.modules["oracle"].set_mutation("none")
        except Exception:
            pass
