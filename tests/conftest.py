import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def beams_mod():
    from pilot_amd import beams as B

    B.lib()
    return B
