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


def sampled_hausdorff(va, fa, vb, fb, k=128):
    """Symmetric Hausdorff distance (index units = cells) between two triangle meshes,
    sampled at every vertex and triangle centroid of each mesh, each sample's distance
    to the other mesh exact (point-triangle distance) over the k triangles with the
    nearest centroids.  A marching-tetrahedra triangle lies in one unit cube (diameter
    <= sqrt(3)) and no point of a triangle is farther than 2/3 of its diameter from its
    centroid, so a triangle whose centroid is c away from a sample is at least
    c - 2/sqrt(3) away: the check below proves that no other triangle is closer."""
    import math
    from scipy.spatial import cKDTree
    from oracle.stencils import point_triangle_distance2

    def one_sided(v, f, w, g):
        p = np.concatenate([v, v[f].mean(axis=1)])
        tri = w[g]
        c, cols = cKDTree(tri.mean(axis=1)).query(p, k=min(k, len(tri)))
        t = tri[cols]
        d2 = point_triangle_distance2(p[:, None, :], t[..., 0, :], t[..., 1, :], t[..., 2, :])
        best = np.sqrt(d2.min(axis=1))
        assert (c[:, -1] - 2.0 / math.sqrt(3.0) >= best).all() or len(tri) <= k
        return float(best.max())

    va, vb = va.astype(np.float64), vb.astype(np.float64)
    return max(one_sided(va, fa, vb, fb), one_sided(vb, fb, va, fa))
