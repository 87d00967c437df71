"""Mesh SDF on the GPU: analytic KATs of the reference (tests/test_geometry.py:53-95),
brute force over all triangles, and the oracle's field."""

import numpy as np
import pytest
import torch

from oracle import sdf as O
from oracle import stencils as S

pytestmark = pytest.mark.gpu


def _box_sdf(P, size=1.0):
    q = np.abs(P) - size / 2
    return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0)


def _nodes(g):
    return np.stack(np.meshgrid(*g.axes(), indexing="ij"), -1)


@pytest.mark.parametrize("h", [0.05, 0.0137])
def test_cube_sdf_exact(device, h):
    from svierosion_amd import shapes
    from svierosion_amd.grid import Grid
    from svierosion_amd.sdf import mesh_to_sdf
    m = shapes.cube(1.0)
    g = Grid.around(*m.bounds, spacing=h, padding=0.3)
    phi = mesh_to_sdf(m, g, band=6).cpu().numpy()
    ex = np.clip(_box_sdf(_nodes(g)), -6 * h, 6 * h)
    err = np.abs(phi - ex).max() / h
    print(f"cube SDF, h={h}: max error {err:.2e} h", flush=True)
    assert err < 2e-5  # reference bound
    assert err < 1e-6  # measured 5.3e-7 / 5.9e-7 h: the float32 rounding of phi itself


def test_torus_band_vs_brute_force_and_sign(device):
    from svierosion_amd import shapes
    from svierosion_amd.grid import Grid
    from svierosion_amd.sdf import mesh_to_sdf
    m = shapes.torus(0.6, 0.25, 48, 24)
    h = 0.03
    g = Grid.around(*m.bounds, spacing=h, padding=0.25)
    phi = mesh_to_sdf(m, g, band=6).reshape(-1)
    band = torch.nonzero(phi.abs() < 5.99 * h).squeeze(1)
    sel = band[torch.randperm(band.numel(), generator=torch.Generator().manual_seed(0))[:3000].to(band.device)]
    pts = g.node_coords(sel).cpu().numpy()
    tri = m.triangles()
    best = np.full(len(pts), np.inf)
    for c in range(0, len(tri), 256):
        t = tri[c:c + 256]
        d2 = S.point_triangle_distance2(pts[:, None], t[None, :, 0], t[None, :, 1], t[None, :, 2])
        best = np.minimum(best, d2.min(1))
    exact = np.sqrt(best)
    got = phi[sel].abs().double().cpu().numpy()
    assert np.abs(got - exact).max() < 1e-4 * h      # reference bound 1e-4 h
    rho = np.hypot(pts[:, 0], pts[:, 1])
    rr = np.hypot(rho - 0.6, pts[:, 2])
    far = np.abs(rr - 0.25) > 0.01
    # inside the tube iff the node is inside the polygonal torus; away from the facets the analytic test agrees
    assert np.array_equal((phi[sel].cpu().numpy() < 0)[far], (rr < 0.25)[far])


def test_icosphere_error_is_facet_sag(device):
    from svierosion_amd import shapes
    from svierosion_amd.grid import Grid
    from svierosion_amd.sdf import mesh_to_sdf
    m = shapes.icosphere(4, 0.8)
    sag = 0.8 - np.min(np.linalg.norm(m.triangles().mean(1), axis=1))
    h = 0.04
    g = Grid.around(*m.bounds, spacing=h, padding=0.3)
    phi = mesh_to_sdf(m, g, band=6).cpu().numpy()
    ex = np.clip(np.linalg.norm(_nodes(g), axis=-1) - 0.8, -6 * h, 6 * h)
    assert np.abs(phi - ex).max() < 1.05 * sag


def test_sdf_matches_oracle_fp64(device):
    """Unclamped band distances vs the oracle run in FP64 (1e-6 h) and the sign field exactly."""
    from svierosion_amd import shapes
    from svierosion_amd.grid import Grid
    from svierosion_amd.sdf import mesh_to_sdf
    m = shapes.sh_fragment(level=3, radius=1.0, seed=3)
    h = 0.04
    g = Grid.around(*m.bounds, spacing=h, padding=0.5)
    phi, dist, tid = mesh_to_sdf(m, g, band=6, return_aux=True)
    ophi, odist, otid = O.mesh_to_sdf(m.vertices, m.faces, g.origin, h, g.shape, dtype=np.float64)
    d = dist.cpu().numpy()
    od = odist.astype(np.float64)
    within = od < 6.0
    err = np.abs(d[within] - od[within]).max()
    print(f"SDF vs FP64 oracle: max distance error {err:.2e} h over {within.sum()} nodes", flush=True)
    assert err < 1e-6  # cells (= h); FP64 oracle, float32 local frames; measured 4.8e-7
    p = phi.cpu().numpy().reshape(-1)
    assert np.array_equal(p < 0, ophi.reshape(-1) < 0)
    assert np.abs(p - ophi.reshape(-1)).max() < 1e-5 * h
