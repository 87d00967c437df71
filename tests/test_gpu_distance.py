"""Point-mesh distance on the device (svimx_point_mesh_distance) against an FP64
brute force over every triangle: the depth annotation of svierosion/api.py:104-194."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _brute_force(points, tri):
    """FP64 min over all triangles of the point-triangle distance (oracle formula)."""
    from oracle.stencils import point_triangle_distance2
    out = np.full(len(points), np.inf)
    for s in range(0, len(tri), 256):
        t = tri[s:s + 256]
        d2 = point_triangle_distance2(points[:, None, :], t[None, :, 0], t[None, :, 1], t[None, :, 2])
        out = np.minimum(out, d2.min(1))
    return np.sqrt(out)


def test_point_mesh_distance_matches_brute_force(device):
    from svierosion_amd.api import point_mesh_distance
    from svierosion_amd.shapes import sh_fragment, torus
    rng = np.random.default_rng(0)
    for mesh in (sh_fragment(level=3, radius=1.0, seed=2), torus(1.0, 0.35, 48, 24)):
        lo, hi = mesh.bounds
        pts = rng.uniform(lo - 0.3, hi + 0.3, size=(3000, 3))
        pts[:500] = mesh.vertices[rng.integers(0, len(mesh.vertices), 500)]  # on the surface
        got = point_mesh_distance(pts, mesh, device="cuda")
        want = _brute_force(pts, mesh.vertices[mesh.faces].astype(np.float64))
        scale = float((hi - lo).max())
        assert np.abs(got - want).max() < 2e-6 * scale
        assert np.abs(got[:500]).max() < 2e-6 * scale


def test_depth_annotation_on_the_job_grid(device):
    """Depth of eroded-surface vertices: device ring search on the job grid = brute force."""
    from svierosion_amd import ErosionParams, erode
    from svierosion_amd.shapes import icosphere
    m = icosphere(3, 0.8)
    res = erode(m, ErosionParams(law="svi", radius=0.2, k=1, s=0, resolution=48, target_fraction=0.9))
    v = res.mesh.vertices
    want = _brute_force(v, m.vertices[m.faces].astype(np.float64))
    got = res.mesh.vertex_data["depth"].astype(np.float64)
    assert np.abs(got - want).max() < 1e-5
    assert got.max() > 0.01
