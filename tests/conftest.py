"""Shared test setup: import paths, the `gpu` marker and the checker build."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "spectral-hohenberg-halperin-dynamics-2d_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    # The oracle is the checker; build it when absent (g++ only, seconds).
    if not os.path.exists(os.path.join(ORACLE, "liboracle.so")):
        subprocess.run(["make", "-C", ORACLE], check=True, stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def gpu_available():
    import shhd  # noqa: F401
    return True
