import os
import sys
import time

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libqmx.so")


@pytest.hookimpl(hookwrapper=True)
def pytest_runtest_makereport(item, call):
    # Report the duration of passed tests on their line (as the reference's
    # conftest does, src/qimpy/conftest.py:7-13).
    outcome = yield
    report = outcome.get_result()
    if report.when == "call" and report.passed:
        report.longrepr = None
        report.duration_str = f"{call.duration:.2f}s"


@pytest.fixture(scope="session")
def gpu():
    """Initialise the device once; skip nothing — GPU tests need a GPU."""
    import torch

    from qimpy_amd import rc

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a ROCm device")
    rc.init()
    print(f"[gpu] {torch.cuda.get_device_name(0)}", flush=True)
    return rc.device
