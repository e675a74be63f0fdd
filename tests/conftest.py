import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    import scp3_amd

    made = []

    def make(device=0):
        c = scp3_amd.Context(device)
        made.append(c)
        return c

    yield make
    for c in made:
        c.close()


@pytest.fixture(scope="session")
def dimer():
    from scp3_amd import system
    return system.prepare(system.input_path("dimer_qtip4pf"), nmodes2=12, nmodes3=0)


@pytest.fixture(scope="session")
def hexamer():
    from scp3_amd import system
    return system.prepare(system.input_path("hexamer_qtip4pf"), nmodes2=48, nmodes3=0)
