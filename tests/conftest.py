import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SCENES = os.path.join(GOLDEN, "scenes")
GOLDEN_SCENES = ("box", "film", "mesh", "room")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def load_golden_scene(name):
    """Parses tests/golden/scenes/<name>.lrd (mesh paths are relative to
    that directory)."""
    import lambda_amd
    cwd = os.getcwd()
    os.chdir(SCENES)
    try:
        return lambda_amd.Scene.load(name + ".lrd")
    finally:
        os.chdir(cwd)


def golden(name):
    path = os.path.join(GOLDEN, name)
    if path.endswith(".npy"):
        return np.load(path)
    return dict(np.load(path))


@pytest.fixture(scope="session")
def scenes():
    return {n: load_golden_scene(n) for n in GOLDEN_SCENES}
