import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libjet_hip.so kernels)")


@pytest.fixture(scope="session")
def ctx():
    from jet_amd import _native
    return _native.default_context()


@pytest.fixture(scope="session")
def golden_permuter():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "permuter_fixtures.json")) as f:
        return json.load(f)


@pytest.fixture
def builtin_rule():
    """Plans created in the test take the library's built-in kernel rule (the
    autotuned table cleared): for tests written to exercise one particular
    kernel.  The compiled-in table comes back afterwards."""
    from jet_amd import _native
    _native.check(_native.lib().jet_hip_tuning_set(0, None))
    yield
    _native.check(_native.lib().jet_hip_tuning_set(-1, None))
