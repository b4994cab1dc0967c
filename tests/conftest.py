import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "friendsrender-r-public_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def box_scene():
    import scenes
    return scenes.box_scene()


@pytest.fixture(scope="session")
def full_scene_small():
    """Full-material scene with a reduced displaced mesh (freq 20 -> 8000 tris)."""
    import scenes
    return scenes.full_scene(freq=20, tex_size=128)
