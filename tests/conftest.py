import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU machine)")


@pytest.fixture(scope="session")
def smd():
    import gpusmoldyn_amd
    gpusmoldyn_amd.lib()
    return gpusmoldyn_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(smd):
    n = smd.device_count()
    assert n >= 1, "no GPU visible: GPU tests must run on the GPU machine"
    return 0
