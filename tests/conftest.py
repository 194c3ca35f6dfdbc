import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def ctx():
    import pico_amd as pa
    c = pa.Context(0)
    print(f"[gpu] {c.device_name}", flush=True)
    yield c
    c.close()
