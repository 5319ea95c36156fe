import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def capi():
    from acoustic_reliefs_amd import capi as C
    C.lib()
    return C


@pytest.fixture(scope="session")
def ctx(capi):
    c = capi.Context(0)
    yield c


@pytest.fixture(scope="session")
def sphere_mesh():
    from acoustic_reliefs_amd import meshes
    return meshes.icosphere(3, 0.5)


@pytest.fixture(scope="session")
def small_relief():
    """A 0.3 m relief with a 16 x 16 top (~700 triangles): fast for the oracle."""
    from acoustic_reliefs_amd import meshes
    h = meshes.smooth_heights(16, seed=3, sigma=2.0, hmax=0.05)
    P, E, top = meshes.relief_box(width=0.3, depth=0.3, base=0.05, grid=16, heights=h, wall_rows=2, bottom_grid=5)
    return P, E, top


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))
