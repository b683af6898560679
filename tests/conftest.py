import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "online-physics-cuda_amd"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI on the card)")


def pytest_report_header(config):
    return "field-hip tests: -m 'not gpu' on the CPU, -m gpu on the card"


@pytest.fixture(scope="session")
def device():
    import field_hip

    return field_hip.require_device()
