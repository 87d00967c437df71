"""Oracle (test infrastructure only): grid sizing and the per-node stencils.

NumPy restatement of the reference's grid arithmetic and numerical kernels.
Arrays are flat C-order grids (axes x, y, z, strides (ny*nz, nz, 1)); ``idx``
holds flat node indices; points are in index (cell) coordinates.
"""

from __future__ import annotations

import math

import numpy as np


# ------------------------------------------------------------------ grid
def next_fast_size(n: int) -> int:
    """Smallest m >= n with prime factors in {2,3,5,7}.  svierosion/grid.py:18-28."""
    m = max(int(n), 1)
    while True:
        r = m
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 1


def grid_around(lo, hi, spacing: float, padding: float):
    """(origin, spacing, shape) of the padded, centred, FFT-sized grid.

    svierosion/grid.py:44-56 (Grid.around with fast=True).
    """
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    need = np.ceil((hi - lo + 2 * padding) / spacing).astype(int) + 1
    shape = tuple(next_fast_size(int(n)) for n in need)
    centre = 0.5 * (lo + hi)
    origin = centre - 0.5 * (np.asarray(shape) - 1) * spacing
    return tuple(float(o) for o in origin), float(spacing), shape


def make_grid(lo, hi, resolution=200, spacing=None, band=6.0, radius=None, law="svi", margin=4):
    """svierosion/api.py:94-101."""
    ext = float((np.asarray(hi) - np.asarray(lo)).max())
    h = spacing if spacing else ext / resolution
    pad = (band + margin + 2) * h
    if law == "svi":
        pad = max(pad, radius + 3 * h)
    return grid_around(lo, hi, h, pad)


def strides_of(shape):
    nx, ny, nz = shape
    return (ny * nz, nz, 1)


def unravel(idx, shape):
    nx, ny, nz = shape
    idx = np.asarray(idx, dtype=np.int64)
    return np.stack([idx // (ny * nz), (idx // nz) % ny, idx % nz], 1)


# ------------------------------------------------------------------ WENO5 / Godunov
def weno5(v1, v2, v3, v4, v5):
    """Jiang-Peng WENO5 derivative from five differences.  svierosion/kernels.py:83-97."""
    p1 = v1 / 3 - 7 * v2 / 6 + 11 * v3 / 6
    p2 = -v2 / 6 + 5 * v3 / 6 + v4 / 3
    p3 = v3 / 3 + 5 * v4 / 6 - v5 / 6
    s1 = 13.0 / 12 * (v1 - 2 * v2 + v3) ** 2 + 0.25 * (v1 - 4 * v2 + 3 * v3) ** 2
    s2 = 13.0 / 12 * (v2 - 2 * v3 + v4) ** 2 + 0.25 * (v2 - v4) ** 2
    s3 = 13.0 / 12 * (v3 - 2 * v4 + v5) ** 2 + 0.25 * (3 * v3 - 4 * v4 + v5) ** 2
    big = np.maximum(np.maximum(np.maximum(v1 * v1, v2 * v2), np.maximum(v3 * v3, v4 * v4)), v5 * v5)
    eps = 1e-6 * big + 1e-10
    a1 = 0.1 / (s1 + eps) ** 2
    a2 = 0.6 / (s2 + eps) ** 2
    a3 = 0.3 / (s3 + eps) ** 2
    return (a1 * p1 + a2 * p2 + a3 * p3) / (a1 + a2 + a3)


def godunov_inward(flat, idx, strides, inv_h):
    """Godunov |grad phi| for V >= 0 with WENO5 one-sided differences.

    svierosion/kernels.py:146-169.
    """
    f = flat
    c = f[idx]
    acc = np.zeros_like(c)
    ih = c.dtype.type(inv_h)
    for s in strides:
        m3, m2, m1 = f[idx - 3 * s], f[idx - 2 * s], f[idx - s]
        p1, p2, p3 = f[idx + s], f[idx + 2 * s], f[idx + 3 * s]
        d0 = (m2 - m3) * ih
        d1 = (m1 - m2) * ih
        d2 = (c - m1) * ih
        d3 = (p1 - c) * ih
        d4 = (p2 - p1) * ih
        d5 = (p3 - p2) * ih
        back = weno5(d0, d1, d2, d3, d4)
        fwd = weno5(d5, d4, d3, d2, d1)
        acc = acc + np.maximum(np.minimum(back, 0) ** 2, np.maximum(fwd, 0) ** 2)
    return np.sqrt(acc)


# ------------------------------------------------------------------ cell fractions
def cube_plane_fraction(s, a, b, c):
    """Volume fraction of the unit cube below a plane at signed distance s.

    a >= b >= c >= 0 are the sorted |normal| components.  Cancellation-free
    piecewise form of the inclusion-exclusion sum: svierosion/kernels.py:104-122,
    derivation report/report.tex:337-365.
    """
    half = 0.5 * (a + b + c)
    y = half - s
    flip = y > half
    y = np.maximum(np.where(flip, 2 * half - y, y), 0)
    tiny = 1e-12
    bs = np.maximum(b, tiny)
    cs = np.maximum(c, tiny)
    q = 3 * y * y - 3 * y * c + c * c
    g1 = y ** 3 / (bs * cs)
    g2 = q / bs
    over_b = np.maximum(y - b, 0)
    g3 = (q - over_b ** 3 / cs) / bs
    g4 = 3 * (2 * y - b - c)
    G = np.where(y <= c, g1, np.where(y <= b, g2, np.where(y <= b + c, g3, g4)))
    over_a = np.maximum(y - a, 0)
    F = G / (6 * a) - over_a ** 3 / (6 * a * bs * cs)
    F = np.where(flip, 1 - F, F)
    return np.clip(F, 0, 1)


def sort3_abs(x, y, z):
    """svierosion/kernels.py:125-131."""
    x, y, z = np.abs(x), np.abs(y), np.abs(z)
    hi = np.maximum(np.maximum(x, y), z)
    lo = np.minimum(np.minimum(x, y), z)
    return hi, x + y + z - hi - lo, lo


def central_gradient(flat, idx, strides):
    f = flat
    return [0.5 * (f[idx + s] - f[idx - s]) for s in strides]


def cell_fractions(flat, idx, strides):
    """Tangent-plane cell fractions.  svierosion/kernels.py:173-181."""
    c = flat[idx]
    gx, gy, gz = central_gradient(flat, idx, strides)
    gn = np.maximum(np.sqrt(gx * gx + gy * gy + gz * gz), 1e-30)
    a, b, cc = sort3_abs(gx / gn, gy / gn, gz / gn)
    return cube_plane_fraction(c / gn, a, b, cc)


# ------------------------------------------------------------------ closest points / interpolation
def closest_points(flat, idx, shape, h, newton_steps=1):
    """Capped Newton projection, clamped to [1, n-3.001].  svierosion/kernels.py:185-208,
    svierosion/levelset.py:129-151: the first step with the central gradient at the
    node, each further step with the tricubic value and gradient at the current point,
    its length capped at 1.25 |value|/h + 0.25 cells."""
    nx, ny, nz = shape
    dt = flat.dtype.type
    c = flat[idx]
    gx, gy, gz = central_gradient(flat, idx, strides_of(shape))
    g2 = np.maximum(gx * gx + gy * gy + gz * gz, 1e-30)
    length = np.abs(c) / np.sqrt(g2)
    cap = 1.25 * np.abs(c) / dt(h) + 0.25
    scale = (c / g2) * np.minimum(cap / np.maximum(length, 1e-30), 1.0)
    ijk = unravel(idx, shape).astype(flat.dtype)
    hi = [dt(n - 3.001) for n in shape]
    p = np.stack([np.minimum(np.maximum(ijk[:, a] - scale * g, 1.0), hi[a])
                  for a, g in enumerate((gx, gy, gz))], 1).astype(flat.dtype)
    for _ in range(newton_steps - 1):  # svierosion/levelset.py:139-145
        val, grad = tricubic_with_gradient(flat, shape, p)
        g2 = np.maximum((grad * grad).sum(1), 1e-30)
        length = np.abs(val) / np.sqrt(g2)
        cap = 1.25 * np.abs(val) / dt(h) + 0.25
        step = (val / g2 * np.minimum(cap / np.maximum(length, 1e-30), 1.0))[:, None] * grad
        p = np.minimum(np.maximum(p - step, 1.0), np.array(hi, dtype=flat.dtype)).astype(flat.dtype)
    return p


def cubic_weights(t):
    """Lagrange cubic weights for nodes -1, 0, 1, 2.  svierosion/kernels.py:134-137."""
    return (-t * (t - 1) * (t - 2) / 6, (t + 1) * (t - 1) * (t - 2) / 2,
            -(t + 1) * t * (t - 2) / 2, (t + 1) * t * (t - 1) / 6)


def tricubic(flat, shape, p):
    """Tensor cubic Lagrange interpolation at index points p (m, 3).  svierosion/kernels.py:212-228."""
    nx, ny, nz = shape
    fl = np.floor(p)
    t = (p - fl).astype(flat.dtype)
    base = fl.astype(np.int64) - 1
    b0 = base[:, 0] * (ny * nz) + base[:, 1] * nz + base[:, 2]
    wx, wy, wz = cubic_weights(t[:, 0]), cubic_weights(t[:, 1]), cubic_weights(t[:, 2])
    out = np.zeros(len(p), dtype=flat.dtype)
    for a in range(4):
        for b in range(4):
            wab = wx[a] * wy[b]
            off = a * ny * nz + b * nz
            for c in range(4):
                out = out + (wab * wz[c]) * flat[b0 + off + c]
    return out


def cubic_dweights(t):
    """Derivatives of the Lagrange cubic weights.  svierosion/kernels.py:140-143."""
    return (-(3 * t * t - 6 * t + 2) / 6, (3 * t * t - 4 * t - 1) / 2,
            -(3 * t * t - 2 * t - 2) / 2, (3 * t * t - 1) / 6)


def tricubic_with_gradient(flat, shape, p):
    """Tricubic value and gradient (per cell) at index points p (m, 3).  svierosion/kernels.py:232-253."""
    nx, ny, nz = shape
    fl = np.floor(p)
    t = (p - fl).astype(flat.dtype)
    base = fl.astype(np.int64) - 1
    b0 = base[:, 0] * (ny * nz) + base[:, 1] * nz + base[:, 2]
    w = [cubic_weights(t[:, a]) for a in range(3)]
    dw = [cubic_dweights(t[:, a]) for a in range(3)]
    val = np.zeros(len(p), dtype=flat.dtype)
    g = [np.zeros(len(p), dtype=flat.dtype) for _ in range(3)]
    for a in range(4):
        for b in range(4):
            off = a * ny * nz + b * nz
            for c in range(4):
                v = flat[b0 + off + c]
                val = val + w[0][a] * w[1][b] * w[2][c] * v
                g[0] = g[0] + dw[0][a] * w[1][b] * w[2][c] * v
                g[1] = g[1] + w[0][a] * dw[1][b] * w[2][c] * v
                g[2] = g[2] + w[0][a] * w[1][b] * dw[2][c] * v
    return val, np.stack(g, 1)


def trilinear(flat, shape, p):
    """svierosion/kernels.py:257-270."""
    nx, ny, nz = shape
    fl = np.floor(p)
    t = (p - fl).astype(flat.dtype)
    base = fl.astype(np.int64)
    b0 = base[:, 0] * (ny * nz) + base[:, 1] * nz + base[:, 2]
    out = np.zeros(len(p), dtype=flat.dtype)
    for a in range(2):
        wa = t[:, 0] if a else 1 - t[:, 0]
        for b in range(2):
            wb = wa * (t[:, 1] if b else 1 - t[:, 1])
            for c in range(2):
                wc = t[:, 2] if c else 1 - t[:, 2]
                out = out + wb * wc * flat[b0 + a * ny * nz + b * nz + c]
    return out


def principal_curvatures(flat, idx, strides, h):
    """k1 >= k2 from second central differences.  svierosion/kernels.py:274-299."""
    f = flat
    dt = f.dtype.type
    h = dt(h)
    sx, sy, sz = strides
    c = f[idx]
    xp, xm, yp, ym, zp, zm = (f[idx + sx], f[idx - sx], f[idx + sy], f[idx - sy],
                              f[idx + sz], f[idx - sz])
    px, py, pz = (xp - xm) / (2 * h), (yp - ym) / (2 * h), (zp - zm) / (2 * h)
    pxx = (xp - 2 * c + xm) / (h * h)
    pyy = (yp - 2 * c + ym) / (h * h)
    pzz = (zp - 2 * c + zm) / (h * h)

    def mixed(a, b):
        return (f[idx + a + b] - f[idx + a - b] - f[idx - a + b] + f[idx - a - b]) / (4 * h * h)

    pxy, pxz, pyz = mixed(sx, sy), mixed(sx, sz), mixed(sy, sz)
    g2 = np.maximum(px * px + py * py + pz * pz, 1e-30)
    gn = np.sqrt(g2)
    ksum = ((pyy + pzz) * px * px + (pxx + pzz) * py * py + (pxx + pyy) * pz * pz
            - 2 * (px * py * pxy + px * pz * pxz + py * pz * pyz)) / (g2 * gn)
    K = (px * px * (pyy * pzz - pyz * pyz) + py * py * (pxx * pzz - pxz * pxz)
         + pz * pz * (pxx * pyy - pxy * pxy)
         + 2 * px * py * (pxz * pyz - pxy * pzz) + 2 * py * pz * (pxy * pxz - pyz * pxx)
         + 2 * px * pz * (pxy * pyz - pxz * pyy)) / (g2 * g2)
    H = 0.5 * ksum
    r = np.sqrt(np.maximum(H * H - K, 0))
    return H + r, H - r


def interface_mask(flat, idx, strides):
    """6-neighbour sign change (or zero).  svierosion/kernels.py:303-309."""
    c = flat[idx]
    hit = np.zeros(len(idx), dtype=bool)
    for s in strides:
        hit |= (flat[idx + s] * c <= 0) | (flat[idx - s] * c <= 0)
    return hit


# ------------------------------------------------------------------ reinitialisation
def minmod(a, b):
    """svierosion/kernels.py:100-101."""
    return np.where(a * b > 0, np.where(np.abs(a) < np.abs(b), a, b), np.zeros_like(a))


def interface_distance(p0, p1, dd, h):
    """Quadratic-ENO subcell distance to the zero towards a neighbour (Min 2010).

    svierosion/kernels.py:312-331.
    """
    dt = p0.dtype.type
    h = dt(h)
    den = p0 - p1
    lin = h * p0 / np.where(den != 0, den, dt(1e-30))
    c2 = 0.5 * dd / (h * h)
    c1 = (p1 - p0) / h
    c0 = 0.5 * (p0 + p1) - c2 * h * h / 4
    disc = c1 * c1 - 4 * c2 * c0
    sq = np.sqrt(np.maximum(disc, 0))
    q = -0.5 * (c1 + np.where(c1 >= 0, sq, -sq))
    xi = c0 / np.where(q != 0, q, dt(1e-30))
    quad = h / 2 + xi
    ok = (disc >= 0) & (np.abs(dd) > 1e-10 * h * h) & (quad >= 0) & (quad <= h)
    return np.minimum(np.maximum(np.where(ok, quad, lin), 0), h)


def reinit_geometry(flat, idx, strides, h, cfl):
    """sgn, local dtau and per-axis subcell crossings.  svierosion/kernels.py:335-359."""
    f = flat
    dt = f.dtype.type
    hh = dt(h)
    c = f[idx]
    sgn = np.sign(c)
    hmin = np.zeros_like(c) + hh
    out = []
    for s in strides:
        m2, m1, p1, p2 = f[idx - 2 * s], f[idx - s], f[idx + s], f[idx + 2 * s]
        d_m1 = c - 2 * m1 + m2
        d_0 = p1 - 2 * c + m1
        d_p1 = p2 - 2 * p1 + c
        cp = (c * p1) < 0
        cm = (c * m1) < 0
        xp = interface_distance(c, p1, minmod(d_0, d_p1), h)
        xm = interface_distance(c, m1, minmod(d_0, d_m1), h)
        hmin = np.minimum(hmin, np.where(cp, xp, hmin))
        hmin = np.minimum(hmin, np.where(cm, xm, hmin))
        out += [cp, cm, np.maximum(xp, dt(1e-6 * h)), np.maximum(xm, dt(1e-6 * h))]
    dtau = dt(cfl) * np.maximum(hmin, dt(1e-3 * h))
    return [sgn, dtau] + out


def reinit_rhs(flat, idx, strides, h, geo):
    """-sgn(phi0)(|grad phi| - 1), ENO2 + subcell fix.  svierosion/kernels.py:363-379."""
    f = flat
    dt = f.dtype.type
    hh = dt(h)
    sgn = geo[0]
    c = f[idx]
    acc = np.zeros_like(c)
    outside = sgn > 0
    for a, s in enumerate(strides):
        cp, cm, xp, xm = geo[2 + 4 * a: 6 + 4 * a]
        m2, m1, p1, p2 = f[idx - 2 * s], f[idx - s], f[idx + s], f[idx + 2 * s]
        mm_p = minmod(p1 - 2 * c + m1, p2 - 2 * p1 + c)
        mm_m = minmod(p1 - 2 * c + m1, c - 2 * m1 + m2)
        Dp = np.where(cp, -c / xp - 0.5 * xp * mm_p / (hh * hh), (p1 - c - 0.5 * mm_p) / hh)
        Dm = np.where(cm, c / xm + 0.5 * xm * mm_m / (hh * hh), (c - m1 + 0.5 * mm_m) / hh)
        pos = np.maximum(np.maximum(Dm, 0) ** 2, np.minimum(Dp, 0) ** 2)
        neg = np.maximum(np.minimum(Dm, 0) ** 2, np.maximum(Dp, 0) ** 2)
        acc = acc + np.where(outside, pos, neg)
    return -sgn * (np.sqrt(acc) - 1.0)


# ------------------------------------------------------------------ point-triangle distance
def point_triangle_distance2(p, a, b, c):
    """Squared point-triangle distance, Ericson sec. 5.1.5 region classification.

    Regions applied in reverse priority so higher priority overrides, as in
    svierosion/kernels.py:382-427.  All inputs (..., 3).
    """
    ab, ac = b - a, c - a
    ap, bp, cp = p - a, p - b, p - c
    d1 = (ab * ap).sum(-1)
    d2 = (ac * ap).sum(-1)
    d3 = (ab * bp).sum(-1)
    d4 = (ac * bp).sum(-1)
    d5 = (ab * cp).sum(-1)
    d6 = (ac * cp).sum(-1)
    va = d3 * d6 - d5 * d4
    vb = d5 * d2 - d1 * d6
    vc = d1 * d4 - d3 * d2
    tiny = 1e-30
    den = va + vb + vc
    den = np.where(np.abs(den) > tiny, den, tiny)
    v = vb / den
    w = vc / den
    e1 = d4 - d3
    e2 = d5 - d6
    r = (va <= 0) & (e1 >= 0) & (e2 >= 0)
    t = e1 / np.maximum(e1 + e2, tiny)
    v = np.where(r, 1 - t, v)
    w = np.where(r, t, w)
    r = (vb <= 0) & (d2 >= 0) & (d6 <= 0)
    t = d2 / np.maximum(d2 - d6, tiny)
    v = np.where(r, 0, v)
    w = np.where(r, t, w)
    r = (d6 >= 0) & (d5 <= d6)
    v = np.where(r, 0, v)
    w = np.where(r, 1, w)
    r = (vc <= 0) & (d1 >= 0) & (d3 <= 0)
    t = d1 / np.maximum(d1 - d3, tiny)
    v = np.where(r, t, v)
    w = np.where(r, 0, w)
    r = (d3 >= 0) & (d4 <= d3)
    v = np.where(r, 1, v)
    w = np.where(r, 0, w)
    r = (d1 <= 0) & (d2 <= 0)
    v = np.where(r, 0, v)
    w = np.where(r, 0, w)
    q = a + v[..., None] * ab + w[..., None] * ac
    return ((p - q) ** 2).sum(-1)


__all__ = [n for n in dir() if not n.startswith("_") and n not in ("math", "np", "annotations")]
