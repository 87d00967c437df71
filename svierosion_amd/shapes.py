"""Seeded test solids for the erosion workloads.

* ``box`` / ``cube``       -- axis-aligned box as 12 outward triangles;
* ``icosphere``            -- subdivided icosahedron projected to a sphere
                              (20 * 4**s triangles);
* ``torus``                -- quad-strip torus split into triangles;
* ``sh_fragment``          -- icosphere with a radial displacement by a sum of
                              real spherical harmonics (the 81,920-triangle
                              fragment at level 6 and the batched mix);
* ``noisy_ellipsoid``      -- icosphere scaled to an ellipsoid with a radial
                              multi-octave value-noise displacement (the
                              1.3M-triangle scan at level 8).

All generators are deterministic given their seed (numpy default_rng).
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
