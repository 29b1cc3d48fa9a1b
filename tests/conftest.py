import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs through libstardust_hip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.load()
    return O


@pytest.fixture(scope="session")
def sdlib():
    from stardust_amd import _lib
    _lib.load()
    return _lib


@pytest.fixture(scope="session")
def warm_model():
    """A warm-start network (GPU) trained for a few epochs on the plain GJK
    directions of eight seeded hull scenes (3,000 hulls of the
    configuration-4 kind in a 0.3 m box each)."""
    from stardust_amd import _lib
    from stardust_amd import hulls as H
    from stardust_amd import warmstart as WS
    _lib.require_gpu()
    data = WS.PairData(80000)
    for k in range(8):
        data.append_scene(H.generate_hulls(0x4B10 + k, 3000, lo=(0.0, 0.0, 0.0), hi=(0.3, 0.3, 0.3)))
    m = WS.WarmStartModel()
    m.train(data, epochs=4, batch=1024, lr=3e-3)
    return m
