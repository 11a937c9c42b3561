import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def orc():
    import clibs
    return clibs.oracle()


@pytest.fixture(scope="session")
def ref():
    import clibs
    r = clibs.ref()
    if r is None:
        pytest.skip("reference build oracle/_ref absent (built only where the reference sources are)")
    return r


@pytest.fixture(scope="session")
def fy():
    import fysics_amd
    return fysics_amd.lib()
