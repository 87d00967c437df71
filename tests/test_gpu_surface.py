"""Device marching tetrahedra (svimx_extract_surface) against the NumPy oracle.

The reference meshes with Lewiner marching cubes (svierosion/surface.py:16-56);
its tests check watertightness, orientation and volume of extracted meshes
(tests/test_sdf.py, tests/test_flows.py).  The device mesher is compared
face for face with oracle.surface.marching_tetrahedra (same algorithm in
FP64) and, after projection, against closed-form volumes.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _levelset(phi, h=1.0, origin=(0.0, 0.0, 0.0), width=3.0):
    import torch
    from svierosion_amd.grid import Grid
    from svierosion_amd.levelset import LevelSet
    g = Grid(tuple(origin), float(h), tuple(phi.shape))
    return LevelSet(torch.as_tensor(phi, dtype=torch.float32), g, width=width)


def _canonical_faces(f):
    """Oriented triangles as tuples with the smallest vertex first (rotation kept)."""
    r = np.argmin(f, axis=1)
    rows = np.arange(len(f))
    return {(f[i, r[i]], f[i, (r[i] + 1) % 3], f[i, (r[i] + 2) % 3]) for i in rows}


def _compare_with_oracle(phi32, ls):
    from scipy.spatial import cKDTree
    from oracle.surface import marching_tetrahedra
    from svierosion_amd.surface import device_mesh
    v_dev, f_dev = device_mesh(ls.flat, ls.band, ls.nb, ls.dims, 0.0)
    v_ref, f_ref = marching_tetrahedra(phi32)
    assert len(v_dev) == len(v_ref) and len(f_dev) == len(f_ref)
    # vertex correspondence by position (FP32 device interpolation vs FP64)
    d, idx = cKDTree(v_ref).query(v_dev.astype(np.float64))
    assert d.max() < 2e-6 * max(phi32.shape)
    assert len(np.unique(idx)) == len(v_ref)
    assert _canonical_faces(idx[f_dev]) == _canonical_faces(f_ref)
    return v_dev, f_dev


def test_ellipsoid_matches_oracle_face_for_face(device):
    n = 56
    ax = np.arange(n) - 27.37
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    phi = (np.sqrt((X / 1.3) ** 2 + Y ** 2 + (Z / 0.8) ** 2) - 14.0).astype(np.float32)
    ls = _levelset(phi)
    _compare_with_oracle(phi, ls)


def test_noisy_field_with_zero_nodes_matches_oracle(device):
    # value-noise perturbation plus exact zeros on nodes: exercises the
    # node-vertex merge (phi == 0) and dropped degenerate triangles.
    rng = np.random.default_rng(7)
    n = 44
    ax = np.arange(n) - 21.5
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    phi = np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 12.0 + rng.uniform(-0.6, 0.6, X.shape)
    phi = np.round(phi * 2) / 2  # many exact zeros
    phi = phi.astype(np.float32)
    ls = _levelset(phi)
    v, f = _compare_with_oracle(phi, ls)
    assert np.count_nonzero(phi == 0) > 100


def test_projected_surface_closed_oriented_accurate(device):
    from svierosion_amd.surface import extract_surface
    h = 0.04
    ax = np.arange(-1.5, 1.5 + 1e-9, h)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    a, b, c = 1.1, 0.9, 0.7
    # distance-like field (exact zero set: the ellipsoid), gradient ~ 1 near it
    r = np.sqrt((X / a) ** 2 + (Y / b) ** 2 + (Z / c) ** 2)
    phi = ((r - 1.0) * 0.9).astype(np.float32)
    ls = _levelset(phi, h=h, origin=(-1.5, -1.5, -1.5))
    m = extract_surface(ls)
    assert m.is_watertight() and m.is_consistently_oriented()
    assert m.signed_volume() > 0
    exact = 4 / 3 * np.pi * a * b * c
    assert abs(m.volume() / exact - 1) < 1.5e-3
    v = m.vertices
    rv = np.sqrt((v[:, 0] / a) ** 2 + (v[:, 1] / b) ** 2 + (v[:, 2] / c) ** 2)
    assert np.abs(rv - 1).max() < 0.02


def test_empty_band_and_no_interface(device):
    from svierosion_amd.surface import extract_surface
    phi = np.full((16, 16, 16), 0.5, dtype=np.float32)
    ls = _levelset(phi, width=0.25)
    assert ls.nb == 0
    assert len(extract_surface(ls).faces) == 0
    phi[6:10, 6:10, 6:10] = 0.2  # band nodes but no sign change
    ls = _levelset(phi, width=0.25)
    assert ls.nb == 64
    assert len(extract_surface(ls).faces) == 0
