import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    lib = os.path.join(ROOT, "oracle", "libcpuref.so")
    if not os.path.exists(lib):  # the oracle is test infrastructure; build it on demand
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU gate: a gpu-marked test on a machine without a GPU fails loudly."""
    import channelflow_amd as cfa
    n = cfa.device_count()
    assert n >= 1, "no HIP device visible"
    return n
