import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (run with -m gpu on the GPU box)")


@pytest.fixture(scope="session")
def room():
    from speedball_amd import scenegen
    return scenegen.room_boxes()


@pytest.fixture(scope="session")
def instanced_small():
    """The instanced-mesh workload at reduced size (same generator, 24k triangles)."""
    from speedball_amd import scenegen
    return scenegen.instanced(res=(8, 4, 6), rays=64, total_tris=24576, n_instances=32, n_lights=4)
