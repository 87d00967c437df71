"""Zero-set extraction as a triangle mesh (API of svierosion/surface.py).

The reference meshes the band with scikit-image's Lewiner marching cubes on
the host (svierosion/surface.py:16-56).  Here the mesh comes from the device
(svimx_extract_surface): marching tetrahedra on the Kuhn (6-tetrahedra)
split of every band cell that straddles the zero set.  The split is
conforming across cells, so the mesh is watertight by construction and has no
ambiguous cases.  Vertices are placed by linear interpolation along
tetrahedron edges and then moved by one capped Newton step onto the zero set
of the tricubic interpolant, the projection of svierosion/surface.py:44-51.
Only the finished vertex and face arrays cross to the host.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi
from .grid import Grid
from .levelset import LevelSet
from .meshio import TriMesh


def project_to_zero_set(field: torch.Tensor, shape, verts: np.ndarray, max_shift: float = 0.5) -> np.ndarray:
    """One capped Newton step of index points onto the zero set of the tricubic interpolant."""
    if len(verts) == 0:
        return verts
    f = field.contiguous().reshape(-1)
    p = torch.as_tensor(verts, dtype=torch.float32, device=f.device).contiguous()
    val = torch.empty(p.shape[0], dtype=torch.float32, device=f.device)
    grad = torch.empty_like(p)
    _abi.check(_abi.lib.svimx_op_tricubic_gradient(_abi.stream(), _abi.ptr(f), *map(int, shape), _abi.ptr(p),
                                                   p.shape[0], _abi.ptr(val), _abi.ptr(grad)))
    g2 = (grad * grad).sum(1).clamp_min(1e-30)
    step = (val / g2).unsqueeze(1) * grad
    n = step.norm(dim=1, keepdim=True)
    step = step * (max_shift / n.clamp_min(max_shift)).clamp(max=1.0)
    return (p - step).double().cpu().numpy()


def extract_surface(ls: LevelSet, project: bool = True, max_shift: float = 0.5) -> TriMesh:
    """Zero set of the level set as an outward-oriented, watertight triangle mesh."""
    if ls.nb == 0:
        return TriMesh(np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int64))
    verts, faces = device_mesh(ls.flat, ls.band32, ls.nb, ls.dims, max_shift if project else 0.0)
    grid = ls.grid
    mesh = TriMesh(np.asarray(grid.origin) + grid.spacing * verts.astype(np.float64), faces.astype(np.int64))
    if len(faces) and mesh.signed_volume() < 0:
        mesh = mesh.flipped()
    return mesh


def device_mesh(phi_flat: torch.Tensor, band: torch.Tensor, nb: int, dims, max_shift: float = 0.0):
    """(vertices in index coordinates float32 (V, 3), faces int32 (F, 3)) of the zero set over ``band``
    (flat node indices of any integer type; the kernels read int32)."""
    if band.dtype != torch.int32 or not band.is_contiguous():
        band = band[: int(nb)].to(dtype=torch.int32).contiguous()
    xyz, faces = C.c_void_p(), C.c_void_p()
    nv, nf = C.c_int64(), C.c_int64()
    st = _abi.stream()
    _abi.check(_abi.lib.svimx_extract_surface(
        st, _abi.ptr(phi_flat), _abi.ptr(band), int(nb), *map(int, dims), float(max_shift),
        C.byref(xyz), C.byref(nv), C.byref(faces), C.byref(nf)), "svimx_extract_surface")
    try:
        v = np.empty((nv.value, 3), dtype=np.float32)
        f = np.empty((nf.value, 3), dtype=np.int32)
        _abi.check(_abi.lib.svimx_copy_to_host(st, v.ctypes.data_as(C.c_void_p), xyz, v.nbytes))
        _abi.check(_abi.lib.svimx_copy_to_host(st, f.ctypes.data_as(C.c_void_p), faces, f.nbytes))
    finally:
        _abi.lib.svimx_free(st, xyz)
        _abi.lib.svimx_free(st, faces)
    return v, f


def sample_field(field: torch.Tensor, grid: Grid, points_world: np.ndarray) -> np.ndarray:
    """Tricubic interpolation of a dense device field at world coordinates."""
    from .levelset import tricubic
    if len(points_world) == 0:
        return np.zeros(0)
    p = torch.as_tensor(grid.to_index(points_world), dtype=torch.float32, device=field.device)
    return tricubic(field.reshape(-1), grid.shape, p).double().cpu().numpy()
