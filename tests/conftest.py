import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libsvimx")


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU")
    return torch.device("cuda")


# Sphere centre shifted off the grid, as in the reference's flow tests (tests/test_flows.py:16-30).
CTR = np.array([0.0123, -0.0311, 0.0217])


def sphere_phi(R, h, pad, distort=False, dtype=np.float64):
    """Signed distance of a sphere on the reference's padded grid (optionally distorted)."""
    from oracle.stencils import grid_around
    origin, h, shape = grid_around([-R] * 3, [R] * 3, h, pad)
    ax = [origin[a] + h * np.arange(shape[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    phi = np.sqrt((X - CTR[0]) ** 2 + (Y - CTR[1]) ** 2 + (Z - CTR[2]) ** 2) - R
    if distort:
        phi = phi * (1 + 0.4 * np.sin(3 * X) * np.cos(2 * Y)) + 0.3 * phi ** 2
    return phi.astype(dtype), origin, h, shape
