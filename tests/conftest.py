import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X) and the built libtm.so")
    config.addinivalue_line("markers", "slow: statistical test taking more than a few seconds")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.build()
    return o


@pytest.fixture(scope="session")
def tm():
    """The HIP integrator.  Fails (never skips) when libtm.so or the GPU is absent."""
    import trophic_maelstorm_amd as tm_
    tm_.lib()
    return tm_
