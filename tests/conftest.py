import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def engine():
    import tesseract_amd as t
    e = t.Engine(0)
    print("engine:", e.describe(), flush=True)
    yield e
    e.close()


@pytest.fixture(scope="session")
def ref():
    from oracle import pyref
    if not pyref.available():
        pytest.skip("oracle/_ref not built (the reference sources are not present here)")
    return pyref
