"""Shared test set-up.

Tests marked ``gpu`` need a HIP device and the built engine library; the rest
run on the CPU (the oracle against golden vectors, the host logic, and that
the C ABI library loads and exports what include/orrery_hip.h declares).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built engine")
    # The oracle is test infrastructure: build it if this checkout has not yet.
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


@pytest.fixture(scope="session")
def ctx():
    import orrery_amd as E
    if E.device_count() == 0:
        pytest.fail("no HIP device visible to a gpu-marked test")
    c = E.Context(0)
    yield c
    c.close()
