import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch first: the product library then binds to torch's HIP runtime (one
# runtime per process), whichever test loads it first.
try:
    import torch  # noqa: F401
except Exception:
    pass

import pytest  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running statistical test")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (built on first use)."""
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The product package; fails (does not skip) when no device is visible."""
    import stocha_amd
    n = stocha_amd.device_count()
    assert n > 0, "GPU tests need a HIP device"
    return stocha_amd
