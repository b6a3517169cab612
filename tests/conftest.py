import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcdk.so on the device)")
    config.addinivalue_line("markers", "slow: long-running")


def pytest_sessionfinish(session, exitstatus):
    """Writes the near-ties the parity checks met to $CDK_NEAR_TIE_LOG."""
    path = os.environ.get("CDK_NEAR_TIE_LOG")
    if not path:
        return
    try:
        from tests.parity import NEAR_TIES
    except Exception:
        return
    import json
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(NEAR_TIES, f, indent=1)
