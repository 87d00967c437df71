"""Marching tetrahedra on the Kuhn split -- CPU checker of svimx_extract_surface.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  The reference meshes the
zero set with scikit-image's Lewiner marching cubes (svierosion/surface.py:
16-56), which is absent from this image; the device path replaces it with
marching tetrahedra, and this module restates that algorithm in NumPy (FP64
edge interpolation, vertices shared by edge key) so the device mesh can be
compared vertex for vertex and face for face.

Pinning: an analytic ellipsoid gives a watertight, consistently oriented mesh
whose enclosed volume matches the closed form to 2e-3 (tests/test_host.py);
a plane through grid nodes exercises the phi == 0 vertex merge.
"""

from __future__ import annotations

import numpy as np

# Kuhn split of the unit cube: corner c = 4*x + 2*y + z; each tetrahedron walks
# from corner 0 to corner 7 along one permutation of the axes.
_AXIS_BIT = (4, 2, 1)
_TETS = []
for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
    c, path = 0, [0]
    for ax in perm:
        c += _AXIS_BIT[ax]
        path.append(c)
    _TETS.append(path)
_TETS = np.asarray(_TETS, dtype=np.int64)           # (6, 4) corner ids
_TET_SIGN = np.array([1, -1, -1, 1, 1, -1], dtype=np.int64)  # parity of each axis permutation
_CORNER = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], dtype=np.int64)


def marching_tetrahedra(phi: np.ndarray):
    """Triangles (vertex positions in index coordinates, faces) of the zero set of phi.

    Nodes with phi < 0 are inside.  A vertex on edge (a inside, b outside) sits at
    a + t (b - a), t = phi_a / (phi_a - phi_b); vertices are shared by edge key, and
    those with t == 1 (phi_b == 0) by node, so coincident points are merged.
    """
    nx, ny, nz = phi.shape
    inside = phi < 0
    # cells with a sign change among their 8 corners
    cnt = np.zeros((nx - 1, ny - 1, nz - 1), dtype=np.int8)
    for c in _CORNER:
        cnt += inside[c[0]:nx - 1 + c[0], c[1]:ny - 1 + c[1], c[2]:nz - 1 + c[2]]
    cells = np.argwhere((cnt > 0) & (cnt < 8))
    if len(cells) == 0:
        return np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int64)
    corner_pos = cells[:, None, :] + _CORNER[None]                       # (C, 8, 3)
    corner_id = (corner_pos[..., 0] * ny + corner_pos[..., 1]) * nz + corner_pos[..., 2]
    corner_val = phi[corner_pos[..., 0], corner_pos[..., 1], corner_pos[..., 2]].astype(np.float64)
    tid = corner_id[:, _TETS].reshape(-1, 4)                             # (6C, 4)
    tval = corner_val[:, _TETS].reshape(-1, 4)
    tpos = corner_pos[:, _TETS].reshape(-1, 4, 3)
    tin = tval < 0
    nin = tin.sum(1)
    verts_key, verts_pos, tris = [], [], []

    def edge_vertices(sel, a_idx, b_idx):
        """Vertex keys/positions on edges (a inside -> b outside) for tets ``sel``."""
        va, vb = tval[sel, a_idx], tval[sel, b_idx]
        t = va / (va - vb)
        ia, ib = tid[sel, a_idx], tid[sel, b_idx]
        pa, pb = tpos[sel, a_idx], tpos[sel, b_idx]
        pos = pa + t[:, None] * (pb - pa)
        at_b = vb == 0
        key_hi = np.where(at_b, ib, np.maximum(ia, ib))
        key_lo = np.where(at_b, -1, np.minimum(ia, ib))
        return np.stack([key_hi, key_lo], 1), pos

    # Orientation is combinatorial: a positively oriented tetrahedron listed as
    # (inside..., outside...) yields outward triangles in the canonical vertex
    # order below; an odd arrangement flips them.
    order = np.argsort(~tin, axis=1, kind="stable")                      # inside corners first
    inv_cnt = np.zeros(len(order), dtype=np.int64)
    for i in range(4):
        for j in range(i + 1, 4):
            inv_cnt += order[:, i] > order[:, j]
    parity = np.repeat(_TET_SIGN[None, :], len(cells), 0).reshape(-1) * np.where(inv_cnt % 2 == 0, 1, -1)
    for k_in in (1, 2, 3):
        sel = np.nonzero(nin == k_in)[0]
        if len(sel) == 0:
            continue
        o = order[sel]
        ins, outs = o[:, :k_in], o[:, k_in:]
        if k_in == 1:
            pairs = [(ins[:, 0], outs[:, q]) for q in range(3)]
            polys = [(0, 1, 2)]
        elif k_in == 3:
            pairs = [(ins[:, q], outs[:, 0]) for q in range(3)]
            polys = [(0, 1, 2)]
        else:
            pairs = [(ins[:, 0], outs[:, 0]), (ins[:, 0], outs[:, 1]), (ins[:, 1], outs[:, 1]),
                     (ins[:, 1], outs[:, 0])]
            polys = [(0, 1, 2), (0, 2, 3)]
        base = sum(len(k) for k in verts_key)
        m = len(sel)
        for a, b in pairs:
            kk, pp = edge_vertices(sel, a, b)
            verts_key.append(kk)
            verts_pos.append(pp)
        v = [base + q * m + np.arange(m) for q in range(len(pairs))]
        flip = parity[sel] < 0
        for a, b, c in polys:
            f = np.stack([v[a], v[b], v[c]], 1)
            f[flip] = f[flip][:, ::-1]
            tris.append(f)
    keys = np.concatenate(verts_key)
    pos = np.concatenate(verts_pos)
    faces = np.concatenate(tris)
    uniq, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    faces = inv[faces]
    good = (faces[:, 0] != faces[:, 1]) & (faces[:, 1] != faces[:, 2]) & (faces[:, 0] != faces[:, 2])
    return pos[first], faces[good]


