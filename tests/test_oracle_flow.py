"""Pins the oracle's speed laws, reinitialisation, simulation loop and inside test
on the CPU, each against something other than the oracle's own formula:

* curvature laws: closed forms for alpha = 1 and 2, finite differences of the
  speed for the diffusivity (alpha < 1 regularised form included), the
  asymptotics of the regularisation (svierosion/flows.py:124-213);
* reinitialisation: the reference's interface-preservation bounds on a
  distorted sphere (tests/test_flows.py:79-88), run on the oracle itself;
* the simulation loop: sphere radii against solve_ivp solutions of the radius
  ODE for three SVI laws (tests/test_flows.py:99-107, same sizes and bounds)
  and four curvature laws (:109-116, at h = 0.1 instead of 0.05), and the exact
  stop at the target volume on the reference's cube case (:120-129);
* inside_mask: analytic inside tests on a box and a torus mesh;
* marching tetrahedra: closed, outward-oriented, ellipsoid volume.
"""

import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle import flow as F
from oracle import sdf as O
from oracle import stencils as S
from tests.conftest import CTR, sphere_phi


def _partials_sum(fn, k1, k2, e=1e-6):
    """dV/dk1 + dV/dk2 by central differences (float64)."""
    return ((fn(k1 + e, k2) - fn(k1 - e, k2)) + (fn(k1, k2 + e) - fn(k1, k2 - e))) / (2 * e)


# ------------------------------------------------------------------ curvature laws
def test_curvature_laws_closed_forms():
    rng = np.random.default_rng(4)
    k1 = rng.uniform(-3, 3, 4000)
    k2 = rng.uniform(-3, 3, 4000)
    H = 0.5 * (k1 + k2)
    V, D = F.CurvatureLaw("mean", 1.0).speed_and_diffusivity(k1, k2)
    assert np.array_equal(V, np.maximum(H, 0))
    assert np.array_equal(D, (H > 0).astype(np.float64))
    V, D = F.CurvatureLaw("mean", 2.0, eps=0.25).speed_and_diffusivity(k1, k2)
    assert np.allclose(V, np.maximum(H, 0) ** 2 + 0.25, rtol=1e-14)
    assert np.allclose(D, 2 * np.maximum(H, 0), rtol=1e-14)
    V, D = F.CurvatureLaw("gauss", 1.0).speed_and_diffusivity(k1, k2)
    convex = (k1 > 0) & (k2 > 0)
    assert np.array_equal(V, np.where(convex, k1 * k2, 0.0))
    assert np.array_equal(D, np.where(convex, k1 + k2, 0.0))


@pytest.mark.parametrize("kind,alpha,delta", [("mean", 0.5, 0.05), ("mean", 1.0, 0.0), ("mean", 1.7, 0.0),
                                              ("gauss", 0.25, 0.01), ("gauss", 1.0, 0.0), ("gauss", 1.5, 0.0)])
def test_curvature_law_diffusivity_is_the_speed_derivative(kind, alpha, delta):
    """D = dV/dk1 + dV/dk2 (svierosion/flows.py:172-173, :203-205) against central
    differences of V, away from the kinks H = 0 and k_i = 0."""
    law = F.CurvatureLaw(kind, alpha, delta=delta)
    rng = np.random.default_rng(8)
    k1 = rng.uniform(-2, 4, 3000)
    k2 = rng.uniform(-2, 4, 3000)
    far = (np.abs(0.5 * (k1 + k2)) > 1e-3) & (np.abs(k1) > 1e-3) & (np.abs(k2) > 1e-3)
    k1, k2 = k1[far], k2[far]
    _, D = law.speed_and_diffusivity(k1, k2)
    num = _partials_sum(lambda a, b: law.speed_and_diffusivity(a, b)[0], k1, k2)
    assert np.abs(D - num).max() < 1e-6 * max(1.0, np.abs(num).max())


def test_regularised_power_asymptotics():
    """x (x + delta)^(alpha-1): x^alpha (1 + O(delta/x)) far from 0, slope delta^(alpha-1) at 0,
    and no offset (svierosion/flows.py:139-150)."""
    for alpha, delta in ((0.5, 1e-3), (0.25, 1e-4)):
        law = F.CurvatureLaw("mean", alpha, delta=delta)
        x = np.array([1e3, 1e4]) * delta
        V, _ = law.speed_and_diffusivity(2 * x, 2 * x)        # H = (k1 + k2) / 2 = 2x
        H = 2 * x
        rel = np.abs(V / H ** alpha - 1)
        assert (rel < 2 * (1 - alpha) * delta / H).all()
        tiny = np.array([1e-9 * delta])
        V0, D0 = law.speed_and_diffusivity(tiny, tiny)
        assert abs(V0[0] / tiny[0] - delta ** (alpha - 1)) < 1e-6 * delta ** (alpha - 1)
        assert abs(D0[0] - delta ** (alpha - 1)) < 1e-6 * delta ** (alpha - 1)
        Vz, _ = law.speed_and_diffusivity(np.zeros(1), np.zeros(1))
        assert Vz[0] == 0.0


def test_curvature_law_rejections():
    for kw in ({"alpha": 0.0}, {"alpha": 0.5}, {"alpha": 1.0, "delta": -1.0}):
        with pytest.raises(ValueError):
            F.CurvatureLaw("mean", **kw)


# ------------------------------------------------------------------ reinitialisation
@pytest.mark.parametrize("h,tol", [(0.08, 0.01), (0.04, 0.005)])
def test_reinitialize_preserves_the_interface(h, tol):
    """Reference tests/test_flows.py:79-88 on the oracle: a distorted sphere field
    reinitialised with 40 iterations is within tol*h of the exact distance near the
    sphere; far from the zero set the field becomes a distance (|grad| ~ 1)."""
    phi, origin, h, shape = sphere_phi(1.0, h, 12 * h, distort=True)
    ls = F.LevelSet(phi.astype(np.float32), shape, h)
    ls.reinitialize(40)
    ax = [origin[a] + h * np.arange(shape[a]) for a in range(3)]
    P = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    x, ph = P[ls.band], ls.flat[ls.band]
    ex = np.linalg.norm(x - CTR, axis=1) - 1.0
    near = np.abs(ex) < 1.5 * h
    err = np.abs(ph - ex)[near].max() / h
    print(f"oracle reinit h={h}: interface error {err:.2e} h (bound {tol})")
    assert err < tol
    # away from the interface, within the 3-node reinitialised shell: distance-like
    shell = (np.abs(ex) > 1.5 * h) & (np.abs(ex) < 3 * h)
    assert np.abs(ph - ex)[shell].max() < 0.2 * h


def test_float32_reinitialisation_error_against_float64():
    """The FP32 reinitialisation (the precision of the HIP path, which equals this FP32
    oracle to 1e-4 h) against the same algorithm in FP64 on the distorted sphere, 20
    iterations: 99.9% of the nodes within 1.5 h of the zero set agree to 1e-5 h and the
    level-set volume to 1e-7 relative.  The few worst nodes (2 of 15,140 above 1e-4 h,
    max 3.5e-4 h) start within ~1e-5 h of zero, where the subcell distance to the
    interface is ill-conditioned."""
    phi64, origin, h, shape = sphere_phi(1.0, 0.05, 0.6, distort=True, dtype=np.float64)
    a = F.LevelSet(phi64.copy(), shape, h, width=6.0)
    b = F.LevelSet(phi64.astype(np.float32), shape, h, width=6.0)
    a.reinitialize(iterations=20, cfl=0.5, dilation=3)
    b.reinitialize(iterations=20, cfl=0.5, dilation=3)
    assert a.flat.dtype == np.float64 and b.flat.dtype == np.float32
    near = np.abs(a.flat) < 1.5 * h
    d = np.abs(b.flat.astype(np.float64) - a.flat)[near] / h
    rel = abs(a.volume() - b.volume()) / a.volume()
    print(f"FP32 vs FP64 reinit: 99.9th pct {np.percentile(d, 99.9):.2e} h, max {d.max():.2e} h, "
          f"{int((d > 1e-4).sum())} of {d.size} above 1e-4 h, volume {rel:.2e} relative")
    assert np.percentile(d, 99.9) < 1e-5
    assert d.max() < 1e-3
    assert rel < 1e-7
    # the HIP path's precision form: the same float32 sweeps with the subcell
    # geometry formed in FP64 (rounded to float32 once): the worst node improves
    c = F.LevelSet(phi64.astype(np.float32), shape, h, width=6.0)
    c.reinitialize(iterations=20, cfl=0.5, dilation=3, geometry64=True)
    d64 = np.abs(c.flat.astype(np.float64) - a.flat)[near] / h
    print(f"FP32 sweeps with FP64 geometry: 99.9th pct {np.percentile(d64, 99.9):.2e} h, max {d64.max():.2e} h")
    assert np.percentile(d64, 99.9) < 1e-5
    assert d64.max() < 1e-4 < d.max()


# ------------------------------------------------------------------ simulation: sphere ODEs
def _interface_radius(ls, origin, shape):
    h = ls.h
    ax = [origin[a] + h * np.arange(shape[a]) for a in range(3)]
    P = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    x, ph = P[ls.band], ls.flat[ls.band]
    near = np.abs(ph) < h
    return np.linalg.norm(x[near] - CTR, axis=1) - ph[near]


@pytest.mark.parametrize("k,s,t,tol", [(1, 0, None, 3e-5), (2, 1, 0.8, 3e-4), (0, 0, None, 3e-5)])
def test_simulation_svi_sphere_radius_ode(k, s, t, tol):
    """Reference tests/test_flows.py:99-107: r = 0.4, r/h = 8, T = 0.15; the radius
    obeys dR/dt = -V(1/2 + 3r/(16R)), integrated by solve_ivp from the closed-form
    speed (not the oracle's law object)."""
    r, T = 0.4, 0.15

    def speed(n):
        v = n ** k
        return v / (t - n) if s else v
    Rex = solve_ivp(lambda _t, y: [-speed(0.5 + 3 * r / (16 * y[0]))], (0, T), [1.0],
                    rtol=1e-12, atol=1e-14).y[0, -1]
    h = r / 8
    phi, origin, h, shape = sphere_phi(1.0, h, r + 10 * h)
    ls = F.LevelSet(phi.astype(np.float32), shape, h)
    reason, steps, time, _, _ = F.Simulation(ls, F.SVILaw(r, k, s, t)).run(target_fraction=None, max_time=T)
    assert reason == "max_time" and abs(time - T) < 1e-12
    Rn = _interface_radius(ls, origin, shape)
    print(f"oracle SVI k={k} s={s}: {steps} steps, |R - R_ode| {abs(Rn.mean() - Rex):.2e}")
    assert abs(Rn.mean() - Rex) < tol
    assert Rn.max() - Rn.min() < 10 * tol + 1e-4


def test_simulation_curvature_sphere_radius_ode():
    """Reference tests/test_flows.py:109-116 at h = 0.1 (T = 0.1, bound 5e-4): mean
    curvature R = sqrt(1 - 2T); Gauss alpha=1 R = (1 - 3T)^(1/3); Gauss alpha=1/4
    R = (1 - 1.5T)^(2/3); and the regularised mean-curvature law alpha = 1/2,
    dR/dt = -(1/R)(1/R + delta)^(-1/2), integrated by solve_ivp."""
    T, h = 0.1, 0.1
    dl = 0.05
    ode = solve_ivp(lambda _t, y: [-(1 / y[0]) * (1 / y[0] + dl) ** -0.5], (0, T), [1.0],
                    rtol=1e-12, atol=1e-14).y[0, -1]
    cases = ((F.CurvatureLaw("mean", 1.0), math.sqrt(1 - 2 * T)),
             (F.CurvatureLaw("gauss", 1.0), (1 - 3 * T) ** (1 / 3)),
             (F.CurvatureLaw("gauss", 0.25, delta=1e-4), (1 - 1.5 * T) ** (2 / 3)),
             (F.CurvatureLaw("mean", 0.5, delta=dl), ode))
    for law, Rex in cases:
        phi, origin, h, shape = sphere_phi(1.0, h, 12 * h)
        ls = F.LevelSet(phi.astype(np.float32), shape, h)
        reason, steps, _, _, _ = F.Simulation(ls, law).run(target_fraction=None, max_time=T)
        assert reason == "max_time"
        Rn = _interface_radius(ls, origin, shape)
        print(f"oracle {law.kind} alpha={law.alpha}: {steps} steps, |R - R_exact| {abs(Rn.mean() - Rex):.2e}")
        assert abs(Rn.mean() - Rex) < 5e-4


# ------------------------------------------------------------------ simulation: exact stop
def _cube_mesh():
    v = np.array([[x, y, z] for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)])
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    f = np.array([t for a, b, c, d in quads for t in ((a, b, c), (a, c, d))])
    t = v[f]
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    flip = (n * t.mean(1)).sum(1) < 0
    f[flip] = f[flip][:, ::-1]
    return v, f


def test_simulation_stops_exactly_at_the_target_volume():
    """Reference tests/test_flows.py:120-129 on the oracle: unit cube, r = 0.2, SVI k=1
    s=1 t=0.9, resolution 60, target 0.93.  The regula falsi lands on 0.93 to five
    places, at least one reinitialisation happens on the way, and the zero set's
    volume (marching tetrahedra, divergence theorem) is 0.93 of the cube within 2e-3."""
    from oracle.surface import marching_tetrahedra
    v, f = _cube_mesh()
    origin, h, shape = S.make_grid(v.min(0), v.max(0), resolution=60, radius=0.2)
    phi, _, _ = O.mesh_to_sdf(v, f, origin, h, shape)
    ls = F.LevelSet(phi.astype(np.float32), shape, h)
    sim = F.Simulation(ls, F.SVILaw(0.2, 1, 1, 0.9))
    # fractions are exact in face cells; the cells on the 12 edges and 8 corners are
    # not planar cuts (measured 8.5e-5 of the volume at h = 1/60)
    assert abs(sim.initial_volume - 1.0) < 2e-4
    reason, steps, _, frac, hist = sim.run(target_fraction=0.93)
    assert reason == "target_volume"
    assert abs(frac - 0.93) < 5e-6
    assert any(r.reinit for r in hist) and not hist[-1].reinit
    assert all(b.volume < a.volume for a, b in zip(hist, hist[1:]))
    vv, ff = marching_tetrahedra(ls.flat.reshape(shape))
    tri = vv[ff].astype(np.float64) * h
    vol = np.einsum("ij,ij->i", tri[:, 0], np.cross(tri[:, 1], tri[:, 2])).sum() / 6.0
    print(f"oracle cube stop: {steps} steps, fraction {frac:.8f}, mesh volume ratio {vol:.5f}")
    assert abs(vol - 0.93) < 2e-3


# ------------------------------------------------------------------ inside test on box and torus
def test_inside_mask_box_and_torus():
    """Analytic inside tests: every node farther than the facet sag from the surface is
    classified as the exact solid classifies it, for signed (winding) and parity modes."""
    from svierosion_amd.shapes import box, torus
    b = box((1.0, 0.6, 0.8), center=(0.05, -0.02, 0.01))
    R0, r0 = 0.6, 0.25
    t = torus(R0, r0, 64, 32)
    sag = R0 * (1 - math.cos(math.pi / 64)) + r0 * (1 - math.cos(math.pi / 32)) + 1e-3
    for mesh, inside, clear in (
            (b, lambda P: (np.abs(P - [0.05, -0.02, 0.01]) < [0.5, 0.3, 0.4]).all(-1),
             lambda P: (np.abs(np.abs(P - [0.05, -0.02, 0.01]) - [0.5, 0.3, 0.4]) > 1e-3).all(-1)),
            (t, lambda P: (np.hypot(P[..., 0], P[..., 1]) - R0) ** 2 + P[..., 2] ** 2 < r0 * r0,
             lambda P: np.abs(np.sqrt((np.hypot(P[..., 0], P[..., 1]) - R0) ** 2 + P[..., 2] ** 2) - r0) > sag)):
        origin, h, shape = S.grid_around(mesh.vertices.min(0), mesh.vertices.max(0), 0.03, 0.2)
        ax = [origin[a] + h * np.arange(shape[a]) for a in range(3)]
        P = np.stack(np.meshgrid(*ax, indexing="ij"), -1)
        ok = clear(P)
        for signed in (True, False):
            ins = O.inside_mask(mesh.vertices, mesh.faces, origin, h, shape, signed=signed)
            assert np.array_equal(ins[ok], inside(P)[ok])


# ------------------------------------------------------------------ marching tetrahedra
def test_marching_tetrahedra_closed_oriented_ellipsoid():
    """oracle.surface.marching_tetrahedra on an analytic ellipsoid field: every edge is
    shared by exactly two triangles with opposite directions (closed, consistently
    oriented), the orientation is outward (positive signed volume), and the volume
    is the ellipsoid's 4/3 pi abc within the linear-interpolation error."""
    from oracle.surface import marching_tetrahedra
    n = 48
    ax = np.arange(n) - 23.41
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    a, b, c = 18.0, 14.0, 11.0
    r = np.sqrt((X / a) ** 2 + (Y / b) ** 2 + (Z / c) ** 2)
    v, f = marching_tetrahedra(((r - 1.0) * 12.0).astype(np.float32))
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    fwd = {tuple(x) for x in e}
    assert len(fwd) == len(e)                                   # no directed edge twice
    assert all((q, p) in fwd for p, q in fwd)                   # each has its reverse
    tri = v[f].astype(np.float64)
    vol = np.einsum("ij,ij->i", tri[:, 0], np.cross(tri[:, 1], tri[:, 2])).sum() / 6.0
    exact = 4.0 / 3.0 * math.pi * a * b * c
    print(f"oracle marching tetrahedra: {len(f)} triangles, volume ratio {vol / exact:.5f}")
    assert abs(vol / exact - 1) < 5e-3


# ------------------------------------------------------------------ N_r on the cube (closed forms)
def cube_surface_points():
    """Face centres, edge midpoints and corners of the unit cube centred at 0, with the
    exterior ball fraction the closed form gives there for r below half the side:
    1/2 (half space), 3/4 (right dihedral), 7/8 (octant) (report/report.tex:124-127)."""
    face = [np.eye(3)[a] * s for a in range(3) for s in (-0.5, 0.5)]
    edge = []
    for a in range(3):
        b, c = [x for x in range(3) if x != a]
        for sb in (-0.5, 0.5):
            for sc in (-0.5, 0.5):
                p = np.zeros(3)
                p[b], p[c] = sb, sc
                edge.append(p)
    corner = [np.array([x, y, z]) for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)]
    return {"face": (np.array(face), 0.5), "edge": (np.array(edge), 0.75), "corner": (np.array(corner), 0.875)}


# discretisation bounds of N_r at res 64, r = 0.1 (R = 6.4 cells): measured 5.8e-6
# (faces), 3.0e-4 (edges), 4.1e-4 (corners); the cells cut by the edges and corners
# are not planar cuts, so their fractions carry the tangent-plane error
CUBE_NR_BOUNDS = {"face": 2e-5, "edge": 1e-3, "corner": 1e-3}


def test_cube_res64_nr_closed_forms():
    """The cube case (side 1, 12 triangles, resolution 64, r = 0.1): the oracle's FFT
    N_r, interpolated tricubically at the face centres, edge midpoints and corners,
    equals 1/2, 3/4 and 7/8 within the discretisation bounds above."""
    v, f = _cube_mesh()
    origin, h, shape = S.make_grid(v.min(0), v.max(0), resolution=64, radius=0.1)
    phi, _, _ = O.mesh_to_sdf(v, f, origin, h, shape)
    ls = F.LevelSet(phi.astype(np.float32), shape, h)
    N = F.exterior_fraction_fft(ls.dense_fractions(), F.ball_kernel_hat(shape, F.ball_cell_weights(0.1, h)), h, 0.1)
    for name, (P, exact) in cube_surface_points().items():
        val = S.tricubic(N.reshape(-1).astype(np.float64), shape, (P - origin) / h)
        err = np.abs(val - exact).max()
        print(f"oracle cube N_r at {len(P)} {name} points: |N - {exact}| <= {err:.2e}")
        assert err < CUBE_NR_BOUNDS[name]


# ------------------------------------------------------------------ the C++ CPU restatement
def test_cpu_port_history_matches_the_numpy_oracle():
    """oracle/cpu/svicpu.cpp (the timed CPU baseline) against the NumPy oracle on the
    sphere, SVI k=1 s=1 t=0.93, 8 steps with two reinitialisations: the same
    reinitialisation steps, fractions within 1e-4 and dt within 1e-3 relative (the
    GPU history bounds); reinitialisation alone within 1e-5 h."""
    import ctypes as C
    from oracle import cpu_port
    r, h = 0.3, 0.05
    phi, origin, h, shape = sphere_phi(1.0, h, r + 10 * h)
    law = F.SVILaw(r, 1, 1, 0.93)
    _, rows, threads = cpu_port.run_steps(phi.astype(np.float32), h, law, 8)
    ref = F.Simulation(F.LevelSet(phi.astype(np.float32), shape, h), law)
    _, _, _, _, hist = ref.run(target_fraction=None, max_steps=8)
    ef = max(abs(vol / ref.initial_volume - b.fraction) for (_, vol, _), b in zip(rows, hist))
    ed = max(abs(dt - b.dt) / b.dt for (dt, _, _), b in zip(rows, hist))
    print(f"C++ port ({threads} threads) vs NumPy oracle over 8 steps: fraction {ef:.2e}, dt {ed:.2e} relative")
    assert [re for _, _, re in rows] == [b.reinit for b in hist] and sum(b.reinit for b in hist) == 2
    assert ef < 1e-4 and ed < 1e-3
    lib = cpu_port._load()
    lib.svicpu_reinit.restype = C.c_int64
    lib.svicpu_reinit.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]
    d, _, h2, shp = sphere_phi(1.0, 0.08, 12 * 0.08, distort=True)
    a = F.LevelSet(d.astype(np.float32), shp, h2)
    a.reinitialize(20)
    b = np.ascontiguousarray(d.astype(np.float32))
    assert lib.svicpu_reinit(b.ctypes.data, *shp, h2, 6.0, 20) == a.band.size
    assert np.abs(b.reshape(-1) - a.flat).max() < 1e-5 * h2
