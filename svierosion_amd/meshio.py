"""Triangle meshes (the ``TriMesh`` type of svierosion/meshio.py:41-131).

``TriMesh`` keeps the reference's fields and behaviour: float64 vertices,
int64 faces, a dict of per-vertex float arrays, validation on construction,
bounds, signed volume by the divergence theorem, and edge topology counts
(boundary / non-manifold / misoriented edges), which choose the sign rule of
the SDF.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class TriMesh:
    vertices: np.ndarray
    faces: np.ndarray
    vertex_data: dict = field(default_factory=dict)

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float64)
        self.faces = np.ascontiguousarray(self.faces, dtype=np.int64)
        if self.vertices.ndim != 2 or self.vertices.shape[1] != 3:
            raise ValueError("vertices must have shape (n, 3)")
        if self.faces.ndim != 2 or self.faces.shape[1] != 3:
            raise ValueError("faces must have shape (m, 3)")

    @property
    def bounds(self) -> np.ndarray:
        return np.stack([self.vertices.min(0), self.vertices.max(0)])

    def triangles(self) -> np.ndarray:
        return self.vertices[self.faces]

    def face_normals(self, normalize: bool = True) -> np.ndarray:
        t = self.triangles()
        n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
        if normalize:
            n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-300)
        return n

    def area(self) -> float:
        return 0.5 * float(np.linalg.norm(self.face_normals(False), axis=1).sum())

    def signed_volume(self) -> float:
        """Enclosed volume, positive for outward orientation (svierosion/meshio.py:80-85):
        sum of (p0 - c) . ((p1 - c) x (p2 - c)) / 6 about the vertex centroid c,
        in FP64 by the library's host loop (svimx_mesh_signed_volume)."""
        import ctypes as C
        from . import _abi
        v = np.ascontiguousarray(self.vertices, dtype=np.float64)
        f = np.ascontiguousarray(self.faces, dtype=np.int64)
        out = C.c_double()
        _abi.check(_abi.lib.svimx_mesh_signed_volume(v.ctypes.data_as(C.c_void_p), len(v), f.ctypes.data_as(C.c_void_p),
                                                     len(f), C.byref(out)), "mesh volume")
        return float(out.value)

    def volume(self) -> float:
        return abs(self.signed_volume())

    def edge_topology(self) -> dict:
        """Edge counts and orientation consistency (svierosion/meshio.py:88-105): edges
        used by one face (boundary), by more than two (non-manifold), and by two
        faces traversing them in the same direction (misoriented); counted by the
        library's host loop (svimx_mesh_edge_topology)."""
        import ctypes as C
        from . import _abi
        f = np.ascontiguousarray(self.faces, dtype=np.int64)
        out = (C.c_int64 * 4)()
        _abi.check(_abi.lib.svimx_mesh_edge_topology(f.ctypes.data_as(C.c_void_p), len(f), len(self.vertices), out),
                   "mesh edge topology")
        return {"edges": int(out[0]), "boundary_edges": int(out[1]), "nonmanifold_edges": int(out[2]),
                "misoriented_edges": int(out[3])}

    def is_watertight(self) -> bool:
        t = self.edge_topology()
        return t["boundary_edges"] == 0 and t["nonmanifold_edges"] == 0

    def is_consistently_oriented(self) -> bool:
        return self.edge_topology()["misoriented_edges"] == 0

    def remove_unreferenced_vertices(self) -> "TriMesh":
        used = np.unique(self.faces)
        remap = -np.ones(len(self.vertices), dtype=np.int64)
        remap[used] = np.arange(len(used))
        return TriMesh(self.vertices[used], remap[self.faces], {k: v[used] for k, v in self.vertex_data.items()})

    def remove_degenerate_faces(self) -> "TriMesh":
        f = self.faces
        keep = (f[:, 0] != f[:, 1]) & (f[:, 1] != f[:, 2]) & (f[:, 0] != f[:, 2])
        return TriMesh(self.vertices, f[keep], dict(self.vertex_data))

    def flipped(self) -> "TriMesh":
        return TriMesh(self.vertices, self.faces[:, ::-1].copy(), dict(self.vertex_data))

    def copy(self) -> "TriMesh":
        return TriMesh(self.vertices.copy(), self.faces.copy(), {k: v.copy() for k, v in self.vertex_data.items()})


def __getattr__(name):
    """``read_ply`` / ``write_ply`` live in ``plyio`` (which imports TriMesh from
    here); they resolve lazily so both import paths of the reference work
    (``from svierosion.meshio import read_ply, write_ply``, svierosion/meshio.py:264,313)."""
    if name in ("read_ply", "write_ply"):
        from . import plyio
        return getattr(plyio, name)
    raise AttributeError(name)
