"""Signed distance of a triangle mesh on the grid (API of svierosion/sdf.py).

Host side (NumPy, FP64, once per job): index coordinates of the faces,
midpoint splitting of long triangles (edges <= 8 cells,
svierosion/sdf.py:54-70), per-triangle local frames and node boxes, and the
vertex -> triangle star CSR.  Device side: ``svimx_sdf_build`` (near field,
propagation, star refinement, FP64 SoS sign, final clamp).
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import _abi
from .grid import Grid, get_device
from .meshio import TriMesh


def split_long_triangles(tri: np.ndarray, max_edge: float) -> np.ndarray:
    """Midpoint-subdivide triangles (n, 3, 3) until every edge is <= max_edge."""
    done = []
    todo = tri
    while len(todo):
        e = np.max(np.stack([np.linalg.norm(todo[:, (q + 1) % 3] - todo[:, q], axis=1) for q in range(3)], 1), 1)
        small = e <= max_edge
        done.append(todo[small])
        big = todo[~small]
        if not len(big):
            break
        a, b, c = big[:, 0], big[:, 1], big[:, 2]
        ab, bc, ca = 0.5 * (a + b), 0.5 * (b + c), 0.5 * (c + a)
        todo = np.concatenate([np.stack(t, 1) for t in ((a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca))])
    return np.concatenate(done) if done else np.zeros((0, 3, 3))


def _point_ids(points: np.ndarray) -> np.ndarray:
    """Ids shared by exactly coincident points (n, 3): unique over the 24-byte
    rows viewed as one opaque key (cheaper than a row-wise lexicographic
    unique; the ids only group coincident points)."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    key = p.view(np.dtype((np.void, 24))).ravel()
    _, inv = np.unique(key, return_inverse=True)
    return inv.reshape(-1)


def _vertex_ids(tri: np.ndarray) -> np.ndarray:
    """Ids of exactly coincident vertices of a triangle soup (n, 3)."""
    return _point_ids(tri.reshape(-1, 3))[: 3 * len(tri)].reshape(-1, 3)


def _prepare_native(mesh: TriMesh, grid: Grid, exact_band: float):
    """The arrays of _prepare from the library's host loop (svimx_sdf_prepare),
    or None when some triangle needs splitting."""
    import ctypes as C
    v = np.ascontiguousarray(mesh.vertices, dtype=np.float64)
    f = np.ascontiguousarray(mesh.faces, dtype=np.int64)
    nf, nv = len(f), len(v)
    tri_idx = np.empty((nf, 3, 3), np.float64)
    org = np.empty((nf, 3), np.int32)
    local = np.empty((nf, 3, 3), np.float32)
    box = np.empty((nf, 6), np.int32)
    vid = np.empty((nf, 3), np.int32)
    star = np.empty(3 * nf, np.int32)
    deg = np.empty(max(nv, 1), np.int32)
    first = np.empty(max(nv, 1), np.int32)
    origin = np.ascontiguousarray(grid.origin, dtype=np.float64)
    shape = np.ascontiguousarray(grid.shape, dtype=np.int32)
    nvid, nsplit = C.c_int64(), C.c_int64()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    _abi.check(_abi.lib.svimx_sdf_prepare(ptr(v), nv, ptr(f), nf, ptr(origin), float(grid.spacing), ptr(shape),
                                          float(exact_band), 8.0, ptr(tri_idx), ptr(org), ptr(local), ptr(box),
                                          ptr(vid), ptr(star), ptr(deg), ptr(first), C.byref(nvid), C.byref(nsplit)),
               "SDF preparation")
    if nsplit.value:
        return None
    k = nvid.value
    return tri_idx, tri_idx, org, local, box, vid, star, deg[:k], first[:k]


def _prepare(mesh: TriMesh, grid: Grid, exact_band: float):
    native = _prepare_native(mesh, grid, exact_band)
    if native is not None:
        return native
    h = grid.spacing
    shape = np.asarray(grid.shape)
    tri_idx = (mesh.vertices[mesh.faces] - np.asarray(grid.origin)) / h
    tri = split_long_triangles(tri_idx, 8.0)
    org = np.floor(tri[:, 0]).astype(np.int64)
    local = (tri - org[:, None, :]).astype(np.float32)
    lo = np.clip(np.ceil(tri.min(1) - exact_band).astype(np.int64), 0, shape - 1)
    hi = np.clip(np.floor(tri.max(1) + exact_band).astype(np.int64), 0, shape - 1)
    box = np.concatenate([lo, hi], 1).astype(np.int32)
    if len(tri) == len(mesh.faces):
        vid = _point_ids(mesh.vertices)[mesh.faces]
    else:
        vid = _vertex_ids(tri)
    flat_v = vid.reshape(-1).astype(np.int64)
    # stable order of the triangle corners by vertex id (one sort of a combined key)
    n_c = len(flat_v)
    order = np.sort(flat_v * n_c + np.arange(n_c)) % n_c if n_c else np.zeros(0, np.int64)
    nv = int(flat_v.max()) + 1 if flat_v.size else 0
    deg = np.bincount(flat_v, minlength=nv)
    first = np.cumsum(deg) - deg
    return tri_idx, tri, org, local, box, vid, order // 3, deg, first


def mesh_to_sdf(mesh: TriMesh, grid: Grid, band: float = 6.0, exact_band: float = 2.0, device=None,
                sign_axes=None, max_pairs: int | None = None, verbose: bool = False,
                return_aux: bool = False, keep_triangles: dict | None = None):
    """Signed distance of ``mesh`` on ``grid`` (negative inside), clamped to +-band*h.

    Returns a dense float32 device tensor of shape ``grid.shape``; with
    ``return_aux`` also the unclamped distance (cells) and closest split triangle.
    ``keep_triangles`` (a dict) receives the device split triangles in local
    frames ("local", "org"), reusable by :func:`triangle_distances`.
    """
    dev = _abi.require_device(get_device(device))
    t0 = time.time()
    topo = mesh.edge_topology()
    closed = topo["boundary_edges"] == 0 and topo["nonmanifold_edges"] == 0
    signed = topo["misoriented_edges"] == 0
    axes = (0,) if closed else (0, 1, 2)
    if sign_axes is not None:
        axes = tuple(sign_axes)
    mask = 0
    for a in axes:
        mask |= 1 << int(a)
    tri_idx, tri, org, local, box, vid, star, deg, first = _prepare(mesh, grid, exact_band)
    t1 = time.time()
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    d_local = up(local.reshape(-1, 9), np.float32)
    d_org = up(org, np.int32)
    d_box = up(box, np.int32)
    d_vid = up(vid, np.int32)
    d_star = up(star, np.int32)
    d_deg = up(deg, np.int32)
    d_first = up(first, np.int32)
    d_faces = up(tri_idx.reshape(-1, 9), np.float64)
    n = grid.size
    phi = torch.empty(grid.shape, dtype=torch.float32, device=dev)
    dist = torch.empty(n, dtype=torch.float32, device=dev) if return_aux else None
    tid = torch.empty(n, dtype=torch.int32, device=dev) if return_aux else None
    ws = _abi.workspace(grid.shape, dev)
    _abi.check(_abi.lib.svimx_sdf_build(_abi.stream(), _abi.ptr(d_local), _abi.ptr(d_org), _abi.ptr(d_box), len(tri),
                                        _abi.ptr(d_vid), _abi.ptr(d_first), _abi.ptr(d_deg), _abi.ptr(d_star),
                                        len(deg), _abi.ptr(d_faces), len(mesh.faces), mask, int(signed),
                                        *map(int, grid.shape), float(grid.spacing), float(band), float(exact_band),
                                        _abi.ptr(phi), _abi.ptr(dist), _abi.ptr(tid), _abi.ptr(ws), ws.numel()),
               "mesh_to_sdf")
    if keep_triangles is not None:
        keep_triangles["local"], keep_triangles["org"] = d_local, d_org
    if verbose:
        print(f"[sdf] {len(mesh.faces)} triangles ({len(tri)} after splitting), {grid.describe()}: "
              f"host prep {t1 - t0:.2f}s, device {time.time() - t1:.2f}s")
    if return_aux:
        return phi, dist, tid
    return phi


def triangle_frames(mesh: TriMesh, grid: Grid, device=None) -> dict:
    """Split triangles of ``mesh`` (edges <= 8 cells) in triangle-local float32
    frames on the device, in index coordinates of ``grid``."""
    dev = _abi.require_device(get_device(device))
    tri_idx = (mesh.vertices[mesh.faces] - np.asarray(grid.origin)) / grid.spacing
    tri = split_long_triangles(tri_idx, 8.0)
    org = np.floor(tri[:, 0]).astype(np.int64)
    local = (tri - org[:, None, :]).astype(np.float32)
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    return {"local": up(local.reshape(-1, 9), np.float32), "org": up(org, np.int32)}


def triangle_distances(triangles: dict, grid: Grid, points_index: torch.Tensor, bucket_cells: float = 4.0):
    """Exact unsigned distance (cells) from index-space points (m, 3) to the
    triangles of :func:`triangle_frames` / ``mesh_to_sdf(keep_triangles=...)``:
    bucketed ring search on the device (``svimx_point_mesh_distance``)."""
    loc, org = triangles["local"], triangles["org"]
    p = points_index.to(loc.device, torch.float32).contiguous()
    out = torch.empty(p.shape[0], dtype=torch.float32, device=loc.device)
    if p.shape[0] == 0:
        return out
    _abi.check(_abi.lib.svimx_point_mesh_distance(_abi.stream(), _abi.ptr(loc), _abi.ptr(org), org.shape[0],
                                                  *map(int, grid.shape), float(bucket_cells), _abi.ptr(p),
                                                  p.shape[0], _abi.ptr(out)), "point-mesh distance")
    return out


def point_triangle_distance2(p: torch.Tensor, a: torch.Tensor, b: torch.Tensor, c: torch.Tensor) -> torch.Tensor:
    """Squared point-triangle distances of matching rows (n, 3) on the device."""
    p, a, b, c = torch.broadcast_tensors(p, a, b, c)
    shape = p.shape[:-1]
    pp = p.reshape(-1, 3).to(torch.float32).contiguous()
    tri = torch.stack([a, b, c], -2).reshape(-1, 9).to(torch.float32).contiguous()
    out = torch.empty(pp.shape[0], dtype=torch.float32, device=pp.device)
    _abi.check(_abi.lib.svimx_op_point_triangle_distance(_abi.stream(), _abi.ptr(pp), _abi.ptr(tri), pp.shape[0],
                                                         _abi.ptr(out)))
    return out.reshape(shape)
