import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libdvc_amd.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.load()
    return O


@pytest.fixture(scope="session")
def ctx():
    import digitalvolumecorrelation_amd as D
    c = D.Context(0)
    yield c
    c.close()
