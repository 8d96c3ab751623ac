import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "msplat-street_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ctx():
    from msplat._abi import Context
    c = Context(0)
    yield c
    c.close()


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def cam_from(d, prefix):
    """Camera dict from golden-vector fields."""
    from msplat.camera import TILE
    cam = {k: d[prefix + k] for k in ("viewmat", "projmat", "cam_pos")}
    for k in ("fx", "fy", "cx", "cy"):
        cam[k] = np.float32(d[prefix + k])
    for k in ("W", "H", "dtu"):
        cam[k] = int(d[prefix + k])
    cam["tiles_x"] = (cam["W"] + TILE - 1) // TILE
    cam["tiles_y"] = (cam["H"] + TILE - 1) // TILE
    return cam
