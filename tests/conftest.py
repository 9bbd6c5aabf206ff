import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(TESTS, "golden")
SCENES = ["three_spheres", "shapes", "cornell_box", "hard_shadows", "soft_shadows", "emissive", "dof", "glass"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP core")


def golden(name):
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="session")
def core():
    """One pcl_ctx on device 0 for the whole GPU session."""
    import pencil_amd as pa
    c = pa.Core(0)
    yield c
    c.close()
