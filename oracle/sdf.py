"""Oracle (test infrastructure only): triangle mesh -> signed distance on a grid.

NumPy restatement of svierosion/sdf.py: exact near field by packed
(distance bits, triangle) keys, ring propagation of closest triangles,
vertex-star refinement, and Simulation-of-Simplicity scan-line sign.
Distances are evaluated in the dtype ``dtype`` (float32 = the reference's
precision in global index coordinates; float64 = the FP64 oracle).
"""

from __future__ import annotations

import math

import numpy as np

from .stencils import point_triangle_distance2, unravel

KEY_EMPTY = np.iinfo(np.int64).max


def edge_topology(faces, n_vertices):
    """Boundary / non-manifold / misoriented edge counts.  svierosion/meshio.py:88-105."""
    f = np.asarray(faces, dtype=np.int64)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    srt = np.sort(e, axis=1)
    n = max(int(n_vertices), 1)
    keys = srt[:, 0] * n + srt[:, 1]
    uniq, inv, counts = np.unique(keys, return_inverse=True, return_counts=True)
    direction = np.where(e[:, 0] < e[:, 1], 1, -1)
    dsum = np.bincount(inv.reshape(-1), weights=direction, minlength=len(uniq))
    manifold = counts == 2
    return {"edges": int(len(uniq)), "boundary_edges": int((counts == 1).sum()),
            "nonmanifold_edges": int((counts > 2).sum()),
            "misoriented_edges": int((manifold & (dsum != 0)).sum())}


def split_long_triangles(tri, max_edge):
    """1->4 midpoint splits until every edge <= max_edge.  svierosion/sdf.py:54-70."""
    out = []
    while len(tri):
        e = np.stack([np.linalg.norm(tri[:, 1] - tri[:, 0], axis=1),
                      np.linalg.norm(tri[:, 2] - tri[:, 1], axis=1),
                      np.linalg.norm(tri[:, 0] - tri[:, 2], axis=1)], 1).max(1)
        ok = e <= max_edge
        out.append(tri[ok])
        big = tri[~ok]
        if not len(big):
            break
        a, b, c = big[:, 0], big[:, 1], big[:, 2]
        ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
        tri = np.concatenate([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1),
                              np.stack([ca, bc, c], 1), np.stack([ab, bc, ca], 1)])
    return np.concatenate(out) if out else np.zeros((0, 3, 3))


def _pack(d, t):
    d = d.astype(np.float32)
    return (d.view(np.int32).astype(np.int64) << 32) | t.astype(np.int64)


def _unpack(keys):
    found = keys != KEY_EMPTY
    dist = np.where(found, (keys >> 32).astype(np.int32).view(np.float32), np.float32(np.inf))
    tid = np.where(found, keys & 0xFFFFFFFF, -1)
    return dist.astype(np.float32), tid.astype(np.int64)


def exact_near_field(tri, shape, exact_band, dtype=np.float32, chunk=1 << 20):
    """Packed keys of all nodes within exact_band cells of a triangle.  svierosion/sdf.py:100-145."""
    shape = np.asarray(shape)
    n = int(np.prod(shape))
    keys = np.full(n, KEY_EMPTY, dtype=np.int64)
    lo = np.clip(np.ceil(tri.min(1) - exact_band).astype(np.int64), 0, shape - 1)
    hi = np.clip(np.floor(tri.max(1) + exact_band).astype(np.int64), 0, shape - 1)
    dims = np.maximum(hi - lo + 1, 0)
    counts = dims.prod(1)
    eb = dtype(exact_band * (1 + 1e-6))
    T = tri.astype(dtype)
    starts = np.concatenate([[0], np.cumsum(counts)])
    t0 = 0
    while t0 < len(tri):
        t1 = int(np.searchsorted(starts, starts[t0] + chunk, side="right")) - 1
        t1 = min(max(t1, t0 + 1), len(tri))
        sel = np.arange(t0, t1)
        cnt = counts[sel]
        t0 = t1
        if cnt.sum() == 0:
            continue
        loc = np.repeat(np.arange(len(sel)), cnt)
        q = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        d_ = dims[sel][loc]
        nyz = d_[:, 1] * d_[:, 2]
        ii = lo[sel][loc, 0] + q // nyz
        rem = q % nyz
        jj = lo[sel][loc, 1] + rem // d_[:, 2]
        kk = lo[sel][loc, 2] + rem % d_[:, 2]
        p = np.stack([ii, jj, kk], 1).astype(dtype)
        t = T[sel[loc]]
        d = np.sqrt(point_triangle_distance2(p, t[:, 0], t[:, 1], t[:, 2]))
        keep = d <= eb
        flat = (ii * shape[1] * shape[2] + jj * shape[2] + kk)[keep]
        np.minimum.at(keys, flat, _pack(d[keep], sel[loc][keep]))
    return keys


OFFSETS26 = [(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1) if (i, j, k) != (0, 0, 0)]


def _dilate26(mask):
    out = mask.copy()
    for a in range(3):
        src = out.copy()
        hi = [slice(None)] * 3
        lo = [slice(None)] * 3
        hi[a] = slice(1, None)
        lo[a] = slice(0, -1)
        out[tuple(hi)] |= src[tuple(lo)]
        out[tuple(lo)] |= src[tuple(hi)]
    return out


def propagate(dist, tid, T, shape, rings, dtype=np.float32):
    """Ring-by-ring closest-triangle propagation (Jacobi per ring).  svierosion/sdf.py:163-202."""
    known = (tid >= 0).reshape(shape)
    sx, sy = shape[1] * shape[2], shape[2]
    offs = [o[0] * sx + o[1] * sy + o[2] for o in OFFSETS26]
    n = dist.size
    prev = None
    for _ in range(rings):
        ring = _dilate26(known) & ~known
        for a in range(3):
            sl = [slice(None)] * 3
            sl[a] = 0
            ring[tuple(sl)] = False
            sl[a] = shape[a] - 1
            ring[tuple(sl)] = False
        upd = np.nonzero(ring.reshape(-1))[0]
        if prev is not None:
            upd = np.concatenate([upd, prev])
        if upd.size == 0:
            break
        p = unravel(upd, shape).astype(dtype)
        best = dist[upd].copy()
        best_t = tid[upd].copy()
        for o in offs:
            nb = np.clip(upd + o, 0, n - 1)
            t = tid[nb]
            ok = t >= 0
            if not ok.any():
                continue
            tt = T[np.maximum(t, 0)]
            d = np.sqrt(point_triangle_distance2(p, tt[:, 0], tt[:, 1], tt[:, 2])).astype(np.float32)
            better = ok & (d < best)
            best = np.where(better, d, best)
            best_t = np.where(better, t, best_t)
        dist[upd] = best
        tid[upd] = best_t
        known |= ring
        prev = np.nonzero(ring.reshape(-1))[0]


def vertex_ids(tri):
    """Ids of exactly coincident vertices in a triangle soup.  svierosion/sdf.py:205-208."""
    _, inv = np.unique(tri.reshape(-1, 3), axis=0, return_inverse=True)
    return inv.reshape(-1)[: 3 * len(tri)].reshape(-1, 3)


def refine_with_stars(dist, tid, T, vid, nodes, shape, passes=2, dtype=np.float32):
    """Re-test the vertex stars of each node's closest triangle.  svierosion/sdf.py:211-265."""
    flat_v = vid.reshape(-1)
    order = np.argsort(flat_v, kind="stable")
    star = order // 3
    deg = np.bincount(flat_v, minlength=int(flat_v.max()) + 1)
    first = np.cumsum(deg) - deg
    for _ in range(passes):
        changed = 0
        t_all = tid[nodes]
        ok = t_all >= 0
        u, t = nodes[ok], t_all[ok]
        if u.size == 0:
            break
        v = vid[t].reshape(-1)
        d = deg[v]
        pair_node = np.repeat(np.repeat(np.arange(u.size), 3), d)
        start = np.repeat(first[v], d)
        offs = np.arange(int(d.sum())) - np.repeat(np.cumsum(d) - d, d)
        cand = star[start + offs]
        p = unravel(u, shape).astype(dtype)[pair_node]
        tt = T[cand]
        dd = np.sqrt(point_triangle_distance2(p, tt[:, 0], tt[:, 1], tt[:, 2])).astype(np.float32)
        best = np.full(u.size, KEY_EMPTY, dtype=np.int64)
        np.minimum.at(best, pair_node, _pack(dd, cand))
        bd, bt = _unpack(best)
        better = bd < dist[u]
        changed += int(better.sum())
        dist[u[better]] = bd[better]
        tid[u[better]] = bt[better]
        if changed == 0:
            break


def orient_sos(x1, y1, x2, y2):
    """Sign of det[(x1,y1),(x2,y2)] with Simulation-of-Simplicity ties.  svierosion/sdf.py:269-282."""
    a = x1 * y2 - y1 * x2
    s = np.sign(a).astype(np.int8)
    zero = s == 0
    if zero.any():
        tie = np.where(y2 > y1, -1, np.where(y2 < y1, 1, np.where(x1 > x2, -1, np.where(x1 < x2, 1, 0))))
        s = np.where(zero, tie.astype(np.int8), s)
    return s, a


def crossings_along_axis(tri_idx, shape, axis, signed):
    """Inside mask from crossings of grid lines parallel to ``axis``.  svierosion/sdf.py:285-340."""
    a, b, c = axis, (axis + 1) % 3, (axis + 2) % 3
    na, nb, nc = shape[a], shape[b], shape[c]
    U, W, Z = tri_idx[:, :, b], tri_idx[:, :, c], tri_idx[:, :, a]
    j0 = np.clip(np.ceil(U.min(1)), 0, nb).astype(np.int64)
    j1 = np.clip(np.floor(U.max(1)), -1, nb - 1).astype(np.int64)
    k0 = np.clip(np.ceil(W.min(1)), 0, nc).astype(np.int64)
    k1 = np.clip(np.floor(W.max(1)), -1, nc - 1).astype(np.int64)
    dj = np.maximum(j1 - j0 + 1, 0)
    dk = np.maximum(k1 - k0 + 1, 0)
    cnt = dj * dk
    tids = np.nonzero(cnt)[0]
    acc = np.zeros((na + 1) * nb * nc, dtype=np.int64)
    n_t = cnt[tids]
    rep = np.repeat(np.arange(len(tids)), n_t)
    q = np.arange(int(n_t.sum())) - np.repeat(np.cumsum(n_t) - n_t, n_t)
    t = tids[rep]
    jj = j0[t] + q // dk[t]
    kk = k0[t] + q % dk[t]
    x1, x2, x3 = U[t, 0] - jj, U[t, 1] - jj, U[t, 2] - jj
    y1, y2, y3 = W[t, 0] - kk, W[t, 1] - kk, W[t, 2] - kk
    sa, A = orient_sos(x2, y2, x3, y3)
    sb, B = orient_sos(x3, y3, x1, y1)
    sc, C = orient_sos(x1, y1, x2, y2)
    hit = (sa != 0) & (sb == sa) & (sc == sa)
    A, B, C, sa, t, jj, kk = A[hit], B[hit], C[hit], sa[hit], t[hit], jj[hit], kk[hit]
    Ssum = A + B + C
    good = Ssum != 0
    z = (A * Z[t, 0] + B * Z[t, 1] + C * Z[t, 2])[good] / Ssum[good]
    i0 = np.clip(np.floor(z).astype(np.int64) + 1, 0, na)
    flat = (i0 * nb + jj[good]) * nc + kk[good]
    w = -sa[good].astype(np.int64) if signed else np.ones(len(flat), dtype=np.int64)
    np.add.at(acc, flat, w)
    acc = acc.reshape(na + 1, nb, nc)
    run = np.cumsum(acc, axis=0)[:na]
    inside = (run != 0) if signed else (run & 1).astype(bool)
    return np.ascontiguousarray(np.transpose(inside, np.argsort([a, b, c])))


def inside_mask(vertices, faces, origin, h, shape, axes=None, signed=None):
    """svierosion/sdf.py:343-361."""
    topo = edge_topology(faces, len(vertices))
    closed = topo["boundary_edges"] == 0 and topo["nonmanifold_edges"] == 0
    if signed is None:
        signed = topo["misoriented_edges"] == 0
    if axes is None:
        axes = (0,) if closed else (0, 1, 2)
    tri_idx = (vertices[faces] - np.asarray(origin)) / h
    votes = None
    for ax in axes:
        m = crossings_along_axis(tri_idx, shape, ax, signed).astype(np.uint8)
        votes = m if votes is None else votes + m
    return votes * 2 > len(axes)


def mesh_to_sdf(vertices, faces, origin, h, shape, band=6.0, exact_band=2.0, dtype=np.float32,
                sign_axes=None):
    """Dense signed distance (negative inside), clamped to +-band*h.  svierosion/sdf.py:365-415."""
    vertices = np.asarray(vertices, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    tri_idx = (vertices[faces] - np.asarray(origin)) / h
    tri = split_long_triangles(tri_idx, 8.0)
    keys = exact_near_field(tri, shape, exact_band, dtype)
    dist, tid = _unpack(keys)
    T = tri.astype(dtype)
    rings = max(int(math.ceil(band - exact_band)) + 1, 0)
    propagate(dist, tid, T, shape, rings, dtype)
    if len(tri) == len(faces):
        _, inv = np.unique(vertices, axis=0, return_inverse=True)
        vid = inv.reshape(-1)[faces]
    else:
        vid = vertex_ids(tri)
    outer = np.nonzero((dist > np.float32(exact_band)) & (dist <= np.float32(band + 1.0)))[0]
    refine_with_stars(dist, tid, T, vid, outer, shape, dtype=dtype)
    inside = inside_mask(vertices, faces, origin, h, shape, axes=sign_axes).reshape(-1)
    d = np.minimum(dist, np.float32(band)) * np.float32(h)
    return np.where(inside, -d, d).reshape(shape).astype(np.float32), dist, tid
