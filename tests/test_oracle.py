"""Pins the CPU oracle against closed forms, brute force and the reference's own
known-answer tests (tests/test_geometry.py, tests/test_flows.py of the reference).

Each check is independent of the oracle's formula: an analytic solution, a
different algorithm (inclusion-exclusion, brute-force sums, segment distances,
Monte-Carlo overlaps), or an invariant a plausible mistake would break.
"""

import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle import flow as F
from oracle import sdf as O
from oracle import stencils as S
from tests.conftest import CTR, sphere_phi


# ------------------------------------------------------------------ grid
def test_grid_kat_cube_res64():
    origin, h, shape = S.make_grid([-0.5] * 3, [0.5] * 3, resolution=64, radius=0.1)
    assert shape == (90, 90, 90)
    assert h == 1 / 64
    assert np.allclose(origin, -0.6953125, atol=0, rtol=0)


def test_next_fast_size():
    smooth = [n for n in range(1, 500) if S.next_fast_size(n) == n]
    for n in smooth:
        r = n
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        assert r == 1
    assert S.next_fast_size(89) == 90 and S.next_fast_size(11) == 12 and S.next_fast_size(97) == 98


# ------------------------------------------------------------------ cell fractions
def test_cube_plane_fraction_vs_inclusion_exclusion():
    """Reference tests/test_geometry.py:111-128: 20k random planes, FP64 classical formula."""
    rng = np.random.default_rng(1)
    n = rng.normal(size=(20000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    s = (rng.random(20000) - 0.5) * 2.0
    a = -np.sort(-np.abs(n), axis=1)
    got = S.cube_plane_fraction(s.astype(np.float32), *(a[:, i].astype(np.float32) for i in range(3)))
    x = -s + a.sum(1) / 2
    tot = np.zeros_like(x)
    for e in np.ndindex(2, 2, 2):
        v = x - e[0] * a[:, 0] - e[1] * a[:, 1] - e[2] * a[:, 2]
        tot += (-1) ** sum(e) * np.maximum(v, 0) ** 3
    ref = np.clip(tot / (6 * a.prod(1)), 0, 1)
    ok = a[:, 2] > 1e-3
    assert np.abs(got.astype(np.float64) - ref)[ok].max() < 2e-5


def test_cube_plane_fraction_degenerate_normals():
    s = np.linspace(-1, 1, 101).astype(np.float32)
    one, zero = np.ones_like(s), np.zeros_like(s)
    assert np.allclose(S.cube_plane_fraction(s, one, zero, zero), np.clip(0.5 - s, 0, 1), atol=1e-6)
    r = np.float32(2 ** -0.5)
    assert np.isfinite(S.cube_plane_fraction(s, r * one, r * one, zero)).all()
    c = np.float32(3 ** -0.5)
    assert abs(float(S.cube_plane_fraction(np.float32([0.0]), *(np.float32([c]),) * 3)[0]) - 0.5) < 1e-6


def test_fraction_monte_carlo():
    """Fractions of random planes against Monte-Carlo counting in the unit cube."""
    rng = np.random.default_rng(7)
    pts = rng.random((200000, 3)) - 0.5
    for _ in range(8):
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        s = rng.uniform(-0.6, 0.6)
        mc = np.mean(pts @ nrm < -s)
        a = -np.sort(-np.abs(nrm))
        f = S.cube_plane_fraction(np.float64(s), a[0], a[1], a[2])
        assert abs(f - mc) < 4e-3


def test_volume_of_ball_converges():
    R, h = 1.0, 0.05
    phi, origin, h, shape = sphere_phi(R, h, 0.5)
    ls = F.LevelSet(phi.astype(np.float32), shape, h)
    vol = ls.volume()
    assert abs(vol / (4 / 3 * math.pi * R ** 3) - 1) < 2e-3


# ------------------------------------------------------------------ WENO / projection / interpolation
def test_godunov_unit_gradient_for_plane_and_sphere():
    shape = (24, 24, 24)
    ax = [np.arange(n, dtype=np.float64) * 0.1 for n in shape]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    nrm = np.array([0.3, -0.5, 0.81]) / np.linalg.norm([0.3, -0.5, 0.81])
    plane = (nrm[0] * X + nrm[1] * Y + nrm[2] * Z - 1.0).reshape(-1)
    idx = np.array([((i * 24) + j) * 24 + k for i in range(4, 20) for j in range(4, 20) for k in range(4, 20)])
    g = S.godunov_inward(plane, idx, S.strides_of(shape), 1 / 0.1)
    assert np.abs(g - 1).max() < 1e-9
    phi, origin, h, shp = sphere_phi(1.0, 0.05, 0.5)
    flat = phi.reshape(-1)
    band = np.nonzero(np.abs(flat) < 3 * h)[0]
    g = S.godunov_inward(flat, band, S.strides_of(shp), 1 / h)
    assert np.abs(g - 1).max() < 2e-3


def test_weno5_polynomial_exactness():
    """For smooth quadratic data every candidate stencil is exact: WENO5 returns the
    third-order upwind derivative, here the exact derivative of a quadratic at x_{i-1/2}."""
    x = np.arange(-3, 3)
    f = 0.7 * x ** 2 - 0.2 * x + 0.1
    d = np.diff(f)
    got = S.weno5(d[0], d[1], d[2], d[3], d[4])
    # backward difference stencil set: derivative value at the node of the differences' centre
    # equals the exact derivative of the quadratic at x = 0 (node between d[2] and d[3]) + O(0)
    assert abs(got - (2 * 0.7 * 0 - 0.2)) < 1e-12


def test_closest_points_project_onto_plane():
    shape = (30, 30, 30)
    ax = [np.arange(n, dtype=np.float64) for n in shape]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    nrm = np.array([0.2, 0.4, -0.89])
    nrm /= np.linalg.norm(nrm)
    h = 0.1
    phi = (h * (nrm[0] * X + nrm[1] * Y + nrm[2] * Z - 3.3)).reshape(-1)
    idx = np.nonzero(np.abs(phi) < 3 * h)[0]
    idx = idx[(S.unravel(idx, shape) > 5).all(1) & (S.unravel(idx, shape) < 24).all(1)]
    p = S.closest_points(phi, idx, shape, h)
    resid = p @ nrm - 3.3
    assert np.abs(resid).max() < 1e-9


def test_tricubic_exact_for_cubics():
    shape = (12, 13, 14)
    ax = [np.arange(n, dtype=np.float64) for n in shape]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    fn = lambda x, y, z: 0.3 * x ** 3 - 0.1 * x * y * z + 0.05 * y ** 2 * z + z ** 3 / 7 - 2 * y + 1
    f = fn(X, Y, Z).reshape(-1)
    rng = np.random.default_rng(3)
    p = rng.uniform(1, 8.9, size=(500, 3))
    got = S.tricubic(f, shape, p)
    assert np.abs(got - fn(p[:, 0], p[:, 1], p[:, 2])).max() < 1e-9
    lin = lambda x, y, z: 0.3 * x - 0.2 * y + 0.7 * z + 2
    g = S.trilinear(lin(X, Y, Z).reshape(-1), shape, p)
    assert np.abs(g - lin(p[:, 0], p[:, 1], p[:, 2])).max() < 1e-9


def test_tricubic_gradient_exact_for_cubics():
    """The tricubic interpolant reproduces polynomials of degree <= 3 per axis: its
    value and its gradient (per cell) equal the analytic ones."""
    shape = (12, 13, 14)
    ax = [np.arange(n, dtype=np.float64) for n in shape]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    fn = lambda x, y, z: 0.3 * x ** 3 - 0.1 * x * y * z + 0.05 * y ** 2 * z + z ** 3 / 7 - 2 * y + 1
    grad = lambda x, y, z: np.stack([0.9 * x ** 2 - 0.1 * y * z, -0.1 * x * z + 0.1 * y * z - 2,
                                     -0.1 * x * y + 0.05 * y ** 2 + 3 * z ** 2 / 7], 1)
    rng = np.random.default_rng(4)
    p = rng.uniform(1, 8.9, size=(500, 3))
    val, g = S.tricubic_with_gradient(fn(X, Y, Z).reshape(-1), shape, p)
    assert np.abs(val - fn(*p.T)).max() < 1e-9
    assert np.abs(g - grad(*p.T)).max() < 1e-9


def test_newton_steps_converge_onto_a_quadric_zero_set():
    """Extra capped Newton steps (svierosion/levelset.py:139-145) on a field that is
    not a distance function, phi = (|x - c|^2 - R^2) / (2R) (its zero set is the
    sphere |x - c| = R): the tricubic interpolant of this quadratic is exact, so each
    step is a Newton step on the exact field and the distance of the points to the
    sphere falls quadratically; on a plane one step is already exact and further
    steps stay there."""
    shape = (40, 40, 40)
    h = 0.05
    c = np.array([19.3, 20.1, 19.7])
    R = 11.0
    ax = [np.arange(n, dtype=np.float64) for n in shape]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    phi = (h * ((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 - R * R) / (2 * R)).reshape(-1)
    idx = np.nonzero(np.abs(phi) < 3 * h)[0]
    errs = []
    for k in (1, 2, 3, 4):
        p = S.closest_points(phi, idx, shape, h, newton_steps=k)
        errs.append(np.abs(np.linalg.norm(p - c, axis=1) - R).max())
    assert errs[0] > 1e-3 and errs[1] < 0.1 * errs[0] and errs[2] < 0.1 * errs[1] and errs[3] < 1e-9, errs
    nrm = np.array([0.2, 0.4, -0.89])
    nrm /= np.linalg.norm(nrm)
    plane = (h * (nrm[0] * X + nrm[1] * Y + nrm[2] * Z - 13.3)).reshape(-1)
    idx = np.nonzero(np.abs(plane) < 3 * h)[0]
    idx = idx[(S.unravel(idx, shape) > 5).all(1) & (S.unravel(idx, shape) < 34).all(1)]
    p = S.closest_points(plane, idx, shape, h, newton_steps=3)
    assert np.abs(p @ nrm - 13.3).max() < 1e-9


def test_principal_curvatures_of_sphere():
    phi, origin, h, shape = sphere_phi(1.0, 0.04, 0.4)
    flat = phi.reshape(-1)
    band = np.nonzero(np.abs(flat) < 0.5 * h)[0]
    k1, k2 = S.principal_curvatures(flat, band, S.strides_of(shape), h)
    r = 1.0 + flat[band]
    assert np.abs(k1 * r - 1).max() < 5e-3 and np.abs(k2 * r - 1).max() < 5e-3


# ------------------------------------------------------------------ point-triangle distance
def _segment_d2(p, a, b):
    ab = b - a
    t = np.clip(((p - a) * ab).sum(-1) / (ab * ab).sum(-1), 0, 1)
    q = a + t[..., None] * ab
    return ((p - q) ** 2).sum(-1)


def test_point_triangle_distance_vs_segments_and_plane():
    """Independent formula: interior -> plane distance, else min of the three edge distances."""
    rng = np.random.default_rng(5)
    a, b, c = rng.normal(size=(3, 4000, 3))
    p = rng.normal(size=(4000, 3)) * 1.5
    got = S.point_triangle_distance2(p, a, b, c)
    n = np.cross(b - a, c - a)
    area2 = (n * n).sum(-1)
    # barycentric of the plane projection
    q = p - ((p - a) * n).sum(-1)[:, None] * n / area2[:, None]
    u = (np.cross(c - b, q - b) * n).sum(-1) / area2
    v = (np.cross(a - c, q - c) * n).sum(-1) / area2
    w = 1 - u - v
    inside = (u >= 0) & (v >= 0) & (w >= 0)
    plane_d2 = ((p - a) * n).sum(-1) ** 2 / area2
    edge_d2 = np.minimum(np.minimum(_segment_d2(p, a, b), _segment_d2(p, b, c)), _segment_d2(p, c, a))
    ref = np.where(inside, plane_d2, edge_d2)
    good = area2 > 1e-6
    assert np.abs(got - ref)[good].max() < 1e-9


# ------------------------------------------------------------------ ball weights and N_r
@pytest.mark.parametrize("ratio", [3.3, 8.0, 12.7])
def test_ball_weights_sum(ratio):
    """Reference tests/test_flows.py:41-44."""
    w = F.ball_cell_weights(1.0, 1.0 / ratio)
    assert abs(w.sum() / ratio ** 3 / (4 * math.pi / 3) - 1.0) < 5e-5


def test_ball_weights_cut_cells_monte_carlo():
    R = 4.7
    w = F.ball_cell_weights(R, 1.0)
    m = (w.shape[0] - 1) // 2
    rng = np.random.default_rng(11)
    cut = np.argwhere((w > 0.05) & (w < 0.95))
    for idx in cut[rng.choice(len(cut), 12, replace=False)]:
        c = idx - m
        pts = c + rng.random((400000, 3)) - 0.5
        mc = np.mean((pts ** 2).sum(1) <= R * R)
        assert abs(mc - w[tuple(idx)]) < 3e-3
    assert w[m, m, m] == 1.0 and w[0, 0, 0] == 0.0


def test_fft_convolution_equals_direct_sum():
    rng = np.random.default_rng(2)
    shape = (30, 28, 27)
    f = np.zeros(shape, np.float32)
    f[8:22, 8:20, 8:19] = rng.random((14, 12, 11)).astype(np.float32)
    h, r = 0.1, 0.35
    w = F.ball_cell_weights(r, h)
    khat = F.ball_kernel_hat(shape, w)
    N = F.exterior_fraction_fft(f, khat, h, r)
    pts = np.stack([rng.integers(6, 22, 40), rng.integers(6, 20, 40), rng.integers(6, 19, 40)], 1)
    direct = F.exterior_fraction_direct(f, w, h, r, pts)
    assert np.abs(N[pts[:, 0], pts[:, 1], pts[:, 2]] - direct).max() < 2e-6


def test_sphere_exterior_fraction_r_over_h_8():
    """Reference tests/test_flows.py:46-56 (r/h = 8, tol 3e-4)."""
    R, r = 1.0, 0.4
    h = r / 8
    phi, origin, h, shape = sphere_phi(R, h, r + 10 * h)
    ls = F.LevelSet(phi.astype(np.float32), shape, h)
    w = F.ball_cell_weights(r, h)
    N = F.exterior_fraction_fft(ls.dense_fractions(), F.ball_kernel_hat(shape, w), h, r)
    pts = ls.closest_points(ls.interface_nodes())
    vals = S.tricubic(N.reshape(-1), shape, pts)
    assert np.abs(vals - F.sphere_exterior_fraction(r, R)).max() < 3e-4


# ------------------------------------------------------------------ laws
def test_svi_law_n_cap_and_derivative():
    law = F.SVILaw(0.3, k=2, s=1, t=0.93)
    flat = law.raw(0.5)
    assert abs(law.raw(law.n_cap) / flat - 100) < 1e-6
    N = np.linspace(0.05, 0.9, 50)
    eps = 1e-6
    num = (law.speed(N + eps) - law.speed(N - eps)) / (2 * eps)
    ok = N + eps < law.n_cap
    assert np.allclose(law.dspeed(N)[ok], num[ok], rtol=1e-5)
    with pytest.raises(ValueError):
        F.SVILaw(0.3, k=1, s=1, t=0.4)


# ------------------------------------------------------------------ SDF
def _cube_mesh():
    v = np.array([[x, y, z] for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)])
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    f = np.array([t for a, b, c, d in quads for t in ((a, b, c), (a, c, d))])
    t = v[f]
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    flip = (n * t.mean(1)).sum(1) < 0
    f[flip] = f[flip][:, ::-1]
    return v, f


def test_cube_sdf_is_exact():
    """Reference tests/test_geometry.py:53-60 (h = 0.05), analytic box SDF, < 2e-5 h."""
    v, f = _cube_mesh()
    h = 0.05
    origin, h, shape = S.grid_around(v.min(0), v.max(0), h, 0.3)
    phi, _, _ = O.mesh_to_sdf(v, f, origin, h, shape)
    ax = [origin[a] + h * np.arange(shape[a]) for a in range(3)]
    P = np.stack(np.meshgrid(*ax, indexing="ij"), -1)
    q = np.abs(P) - 0.5
    ex = np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0)
    assert np.abs(phi - np.clip(ex, -6 * h, 6 * h)).max() < 2e-5 * h


def test_inside_mask_sphere_mesh_parity_and_winding():
    from svierosion_amd.shapes import icosphere
    m = icosphere(3, 0.8)
    origin, h, shape = S.grid_around(m.vertices.min(0), m.vertices.max(0), 0.05, 0.3)
    ax = [origin[a] + h * np.arange(shape[a]) for a in range(3)]
    P = np.stack(np.meshgrid(*ax, indexing="ij"), -1)
    r = np.linalg.norm(P, axis=-1)
    for signed in (True, False):
        ins = O.inside_mask(m.vertices, m.faces, origin, h, shape, signed=signed)
        clear = np.abs(r - 0.79) > 0.02
        assert np.array_equal(ins[clear], (r < 0.79)[clear])


def test_sampled_hausdorff_measures_a_known_offset():
    """The surface-distance checker of the evolution parity tests: marching-tetrahedra
    spheres whose level differs by delta are delta apart (up to the linear
    interpolation along tetrahedron edges), a mesh is 0 from itself, and a mesh
    with one vertex displaced reads at least that displacement's normal part."""
    from oracle.surface import marching_tetrahedra
    from tests.conftest import sampled_hausdorff
    n = 30
    ax = np.arange(n) - 14.3
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    R = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    va, fa = marching_tetrahedra((R - 9.0).astype(np.float32))
    assert sampled_hausdorff(va, fa, va, fa) < 1e-5
    for delta in (0.02, 0.3):
        vb, fb = marching_tetrahedra((R - 9.0 - delta).astype(np.float32))
        d = sampled_hausdorff(va, fa, vb, fb)
        assert abs(d - delta) < 0.1 * delta, (delta, d)
    vc = va.copy()
    i = int(np.argmax(vc[:, 0]))                     # a vertex at the +x pole: normal ~ x
    vc[i, 0] += 0.5
    assert sampled_hausdorff(va, fa, vc, fa) > 0.45
