"""Seeded test solids for the erosion workloads.

* ``box`` / ``cube``       -- axis-aligned box as 12 outward triangles;
* ``triangular_prism``, ``l_block``, ``cylinder``, ``tube_fragment``
                           -- straight prisms over a planar outline whose cap
                              triangulation is given with it;
* ``icosphere``            -- subdivided icosahedron projected to a sphere
                              (20 * 4**s triangles); ``ellipsoid`` scales it;
* ``torus``                -- quad-strip torus split into triangles;
* ``sh_fragment``          -- icosphere with a radial displacement by a sum of
                              real spherical harmonics (the 81,920-triangle
                              fragment at level 6 and the batched mix);
* ``noisy_ellipsoid``      -- icosphere scaled to an ellipsoid with a radial
                              multi-octave value-noise displacement (the
                              1.3M-triangle scan at level 8).

All generators are deterministic given their seed (numpy default_rng).
``SHAPES`` names the toy solids of the ``shapes`` command; ``sphere_sdf`` and
``box_sdf`` are exact signed distances for tests.
"""

from __future__ import annotations

import numpy as np

from .meshio import TriMesh


def _outward(mesh: TriMesh) -> TriMesh:
    return mesh if mesh.signed_volume() >= 0 else mesh.flipped()


def box(size=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0)) -> TriMesh:
    half = np.asarray(size, dtype=np.float64) / 2
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], dtype=np.float64)
    verts = corners * half + np.asarray(center, dtype=np.float64)
    # each face as a quad of corner ids (bit 2 = x, bit 1 = y, bit 0 = z), split along a diagonal
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    faces = []
    ctr = np.asarray(center, dtype=np.float64)
    for a, b, c, d in quads:
        for tri in ((a, b, c), (a, c, d)):
            p = verts[list(tri)]
            n = np.cross(p[1] - p[0], p[2] - p[0])
            faces.append(tri if np.dot(n, p.mean(0) - ctr) > 0 else tri[::-1])
    return TriMesh(verts, np.asarray(faces, dtype=np.int64))


def cube(edge: float = 1.0, center=(0.0, 0.0, 0.0)) -> TriMesh:
    return box((edge, edge, edge), center)


def _prism(outline, cap, height: float) -> TriMesh:
    """Closed prism over a planar outline, |z| <= height/2.

    ``outline`` (n, 2) is the boundary loop and ``cap`` (m, 3) a triangulation
    of it by outline indices; both may come in either orientation, the walls
    and caps are oriented outward here (svierosion/shapes.py:58-78 builds the
    same solids, triangulating the cap by ear clipping)."""
    p = np.asarray(outline, dtype=np.float64)
    cap = np.asarray(cap, dtype=np.int64)
    n = len(p)
    x, y = p[:, 0], p[:, 1]
    if np.dot(x, np.roll(y, -1)) - np.dot(np.roll(x, -1), y) < 0:      # clockwise: reverse the loop
        p = p[::-1]
        cap = n - 1 - cap
    a, b, c = p[cap[:, 0]], p[cap[:, 1]], p[cap[:, 2]]
    ccw = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]) > 0
    cap = np.where(ccw[:, None], cap, cap[:, ::-1])
    z = 0.5 * height
    verts = np.vstack([np.column_stack([p, np.full(n, -z)]), np.column_stack([p, np.full(n, z)])])
    i = np.arange(n)
    j = (i + 1) % n
    faces = np.vstack([cap[:, ::-1], cap + n, np.column_stack([i, j, j + n]), np.column_stack([i, j + n, i + n])])
    return _outward(TriMesh(verts, faces))


def _fan(n: int) -> np.ndarray:
    """Triangulation of a convex n-gon from its first corner."""
    k = np.arange(1, n - 1)
    return np.column_stack([np.zeros(n - 2, dtype=np.int64), k, k + 1])


def triangular_prism(side: float = 1.0, length: float = 1.0) -> TriMesh:
    """Prism along z over an equilateral triangle of edge ``side`` centred on the axis,
    one corner on +y (svierosion/shapes.py:101-105)."""
    ang = np.pi / 2 + np.array([0.0, 2.0, 4.0]) * np.pi / 3
    r = side / np.sqrt(3.0)
    return _prism(np.column_stack([r * np.cos(ang), r * np.sin(ang)]), _fan(3), length)


def l_block(a: float = 1.0, b: float = 0.5, thickness: float = 0.5) -> TriMesh:
    """An a x a square with its (a-b) x (a-b) corner at +x+y removed, extruded to
    ``thickness`` (svierosion/shapes.py:108-116): one re-entrant edge."""
    h = a / 2
    outline = [(-h, -h), (h, -h), (h, b - h), (b - h, b - h), (b - h, h), (-h, h)]
    # two rectangles sharing the corner square at (-h, -h): [0 1 2 3] and [0 3 4 5]
    cap = [(0, 1, 2), (0, 2, 3), (0, 3, 4), (0, 4, 5)]
    return _prism(outline, cap, thickness)


def cylinder(radius: float = 0.5, height: float = 1.0, segments: int = 256) -> TriMesh:
    """Regular ``segments``-gon prism inscribed in the cylinder (svierosion/shapes.py:119-121)."""
    t = 2 * np.pi * np.arange(segments) / segments
    return _prism(np.column_stack([radius * np.cos(t), radius * np.sin(t)]), _fan(segments), height)


def tube_fragment(r_in: float = 0.6, r_out: float = 1.0, angle: float = 2.2,
                  length: float = 1.6, segments: int = 160) -> TriMesh:
    """Thick-walled tube sector r_in <= rho <= r_out, |theta| <= angle/2, |z| <= length/2,
    shifted by -(r_in + r_out)/2 in x (svierosion/shapes.py:124-138): convex outer wall,
    concave inner wall and four break edges.  The cap is the strip of quads between
    the two arcs."""
    t = np.linspace(-angle / 2, angle / 2, segments + 1)
    outer = np.column_stack([r_out * np.cos(t), r_out * np.sin(t)])
    inner = np.column_stack([r_in * np.cos(t), r_in * np.sin(t)])[::-1]
    m = segments + 1
    k = np.arange(segments)
    o0, o1 = k, k + 1                                  # outer arc, increasing angle
    i0, i1 = 2 * m - 1 - k, 2 * m - 2 - k              # inner arc at the same angles
    cap = np.vstack([np.column_stack([o0, o1, i1]), np.column_stack([o0, i1, i0])])
    mesh = _prism(np.vstack([outer, inner]), cap, length)
    mesh.vertices[:, 0] -= 0.5 * (r_in + r_out)
    return mesh


def ellipsoid(radii=(1.0, 0.7, 0.5), subdivisions: int = 5) -> TriMesh:
    """Icosphere scaled by the semi-axes (svierosion/shapes.py:167-170)."""
    m = icosphere(subdivisions, 1.0)
    return _outward(TriMesh(m.vertices * np.asarray(radii, dtype=np.float64), m.faces))


def icosphere(subdivisions: int = 5, radius: float = 1.0) -> TriMesh:
    g = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, g, 0], [1, g, 0], [-1, -g, 0], [1, -g, 0],
                  [0, -1, g], [0, 1, g], [0, -1, -g], [0, 1, -g],
                  [g, 0, -1], [g, 0, 1], [-g, 0, -1], [-g, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    # the 20 faces: every triple of mutually adjacent vertices (edge length 2 in the unscaled solid)
    d = np.linalg.norm(v[:, None] - v[None], axis=2)
    edge = np.isclose(d, d[d > 1e-9].min())
    faces = []
    for i in range(12):
        for j in range(i + 1, 12):
            if not edge[i, j]:
                continue
            for k in range(j + 1, 12):
                if edge[i, k] and edge[j, k]:
                    n = np.cross(v[j] - v[i], v[k] - v[i])
                    faces.append((i, j, k) if np.dot(n, v[i] + v[j] + v[k]) > 0 else (i, k, j))
    f = np.asarray(faces, dtype=np.int64)
    for _ in range(subdivisions):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
        ends = e[first]
        mid = v[ends[:, 0]] + v[ends[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.vstack([v, mid])
        nf = len(f)
        m01, m12, m20 = (inv[:nf] + base, inv[nf:2 * nf] + base, inv[2 * nf:] + base)
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([m01, f[:, 1], m12], 1),
                            np.stack([m20, m12, f[:, 2]], 1), np.stack([m01, m12, m20], 1)])
    return _outward(TriMesh(radius * v, f))


def torus(major: float = 1.0, minor: float = 0.35, n_major: int = 96, n_minor: int = 48) -> TriMesh:
    u = 2 * np.pi * np.arange(n_major) / n_major
    w = 2 * np.pi * np.arange(n_minor) / n_minor
    U, W = np.meshgrid(u, w, indexing="ij")
    pts = np.stack([(major + minor * np.cos(W)) * np.cos(U), (major + minor * np.cos(W)) * np.sin(U),
                    minor * np.sin(W)], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(n_major), np.arange(n_minor), indexing="ij")
    a = i * n_minor + j
    b = ((i + 1) % n_major) * n_minor + j
    c = ((i + 1) % n_major) * n_minor + (j + 1) % n_minor
    dd = i * n_minor + (j + 1) % n_minor
    f = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, dd], -1).reshape(-1, 3)])
    return _outward(TriMesh(pts, f))


def _real_sph_harm(l: int, m: int, dirs: np.ndarray) -> np.ndarray:
    """Real orthonormal spherical harmonic Y_lm at unit vectors ``dirs`` (n, 3)."""
    from scipy.special import sph_harm_y
    theta = np.arccos(np.clip(dirs[:, 2], -1.0, 1.0))
    phi = np.arctan2(dirs[:, 1], dirs[:, 0])
    y = sph_harm_y(l, abs(m), theta, phi)
    if m > 0:
        return np.sqrt(2.0) * (-1) ** m * y.real
    if m < 0:
        return np.sqrt(2.0) * (-1) ** m * y.imag
    return y.real


def sh_fragment(level: int = 6, radius: float = 50.0, terms: int = 16, max_degree: int = 8,
                amplitude: float = 0.08, seed: int = 0) -> TriMesh:
    """Icosphere of ``radius`` displaced radially by A * S / max|S|, A = amplitude * radius,
    S = sum of ``terms`` real Y_lm with l ~ U{1..max_degree}, m ~ U{-l..l}, N(0,1) coefficients."""
    rng = np.random.default_rng(seed)
    base = icosphere(level, 1.0)
    dirs = base.vertices
    S = np.zeros(len(dirs))
    for _ in range(terms):
        l = int(rng.integers(1, max_degree + 1))
        m = int(rng.integers(-l, l + 1))
        S += rng.normal() * _real_sph_harm(l, m, dirs)
    S *= amplitude * radius / np.abs(S).max()
    return _outward(TriMesh(dirs * (radius + S)[:, None], base.faces))


def value_noise(dirs: np.ndarray, rng: np.random.Generator, freq: float, lattice: int = 64) -> np.ndarray:
    """Smoothstep-trilinear value noise of U(-1, 1) lattice values at freq * dirs."""
    table = rng.uniform(-1.0, 1.0, size=(lattice, lattice, lattice))
    x = freq * dirs + lattice / 2.0
    i0 = np.floor(x).astype(np.int64)
    t = x - i0
    s = t * t * (3.0 - 2.0 * t)
    out = np.zeros(len(dirs))
    for dx in (0, 1):
        wx = s[:, 0] if dx else 1 - s[:, 0]
        for dy in (0, 1):
            wy = s[:, 1] if dy else 1 - s[:, 1]
            for dz in (0, 1):
                wz = s[:, 2] if dz else 1 - s[:, 2]
                out += wx * wy * wz * table[(i0[:, 0] + dx) % lattice, (i0[:, 1] + dy) % lattice,
                                            (i0[:, 2] + dz) % lattice]
    return out


def noisy_ellipsoid(level: int = 8, axes=(60.0, 45.0, 30.0), amplitudes=(3.0, 1.5, 0.75, 0.375),
                    base_frequency: float = 2.0, seed: int = 0) -> TriMesh:
    """Icosphere scaled to semi-axes ``axes``; each vertex moved along its radial direction
    by sum_o amplitudes[o] * noise_o(v_hat) with frequency base_frequency * 2^o."""
    rng = np.random.default_rng(seed)
    base = icosphere(level, 1.0)
    dirs = base.vertices
    e = dirs * np.asarray(axes, dtype=np.float64)
    disp = np.zeros(len(dirs))
    for o, amp in enumerate(amplitudes):
        disp += amp * value_noise(dirs, rng, base_frequency * 2 ** o)
    r = np.linalg.norm(e, axis=1)
    return _outward(TriMesh(e * ((r + disp) / r)[:, None], base.faces))


# The toy solids written by ``python -m svierosion_amd shapes`` (names and sizes of
# svierosion/shapes.py:241-252).
SHAPES = {
    "cube": lambda: cube(1.0),
    "box": lambda: box((1.6, 1.0, 0.6)),
    "prism": lambda: triangular_prism(1.2, 1.2),
    "ellipsoid": lambda: ellipsoid((0.8, 0.55, 0.4), 6),
    "sphere": lambda: icosphere(6, 0.6),
    "lblock": lambda: l_block(1.2, 0.6, 0.6),
    "cylinder": lambda: cylinder(0.45, 1.2),
    "torus": lambda: torus(0.6, 0.25, 192, 96),
    "tube_fragment": lambda: tube_fragment(),
}


def sphere_sdf(points, center=(0.0, 0.0, 0.0), radius: float = 1.0) -> np.ndarray:
    """Exact signed distance to a sphere, negative inside (svierosion/shapes.py:255-257)."""
    p = np.asarray(points, dtype=np.float64) - np.asarray(center, dtype=np.float64)
    return np.linalg.norm(p, axis=-1) - radius


def box_sdf(points, size=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Exact signed distance to an axis-aligned box, negative inside (svierosion/shapes.py:260-265):
    outside, the length of the excess over the half sizes; inside, minus the distance
    to the nearest face."""
    q = np.abs(np.asarray(points, dtype=np.float64) - np.asarray(center, dtype=np.float64)) \
        - 0.5 * np.asarray(size, dtype=np.float64)
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(q.max(axis=-1), 0.0)
    return outside + inside
