import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and runs the HIP path")


@pytest.fixture(scope="session")
def ctx():
    import asdmatch_amd as A

    c = A.Context(0)
    yield c
    c.close()
