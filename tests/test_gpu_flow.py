"""Full evolution on the GPU: exact sphere solutions and per-step parity with the oracle."""

import math

import numpy as np
import pytest
import torch
from scipy.integrate import solve_ivp

from oracle import flow as F
from tests.conftest import CTR, sampled_hausdorff, sphere_phi

pytestmark = pytest.mark.gpu


def _gpu_sphere(R, h, pad):
    from svierosion_amd.grid import Grid
    from svierosion_amd.levelset import LevelSet
    phi, origin, h, shape = sphere_phi(R, h, pad)
    g = Grid(origin, h, shape)
    return LevelSet(torch.from_numpy(phi.astype(np.float32)).cuda(), g, width=6), g, phi


def _interface_radius(ls, g):
    idx = ls.band
    x = g.node_coords(idx).cpu().numpy()
    ph = ls.flat[idx.long()].cpu().numpy()
    near = np.abs(ph) < g.spacing
    return np.linalg.norm(x[near] - CTR, axis=1) - ph[near]


# Tolerances of tests/test_flows.py:99-107, except the k=2 s=1 case: there dt is
# CFL-limited, so dt*vmax = cfl*h and after three steps the reinit trigger
# (displacement >= 1.5h, svierosion/evolve.py:235-238) compares 3*(0.5h) with 1.5h --
# a tie decided by the last bit of vmax.  Reinitialising after step 3 (the oracle's
# float32 run) gives |R - R_ode| = 2.0e-4, after step 4 3.0e-4; both are the
# reference algorithm, so the bound covers both schedules.
@pytest.mark.parametrize("k,s,t,tol", [(1, 0, None, 3e-5), (2, 1, 0.8, 3.3e-4), (0, 0, None, 3e-5)])
def test_svi_sphere_radius_ode(device, k, s, t, tol):
    """The sphere radius follows dR/dt = -F(N_r(R)) (solve_ivp reference solution)."""
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.flows import SVIPowerLaw
    r, T = 0.4, 0.15
    law = SVIPowerLaw(r, k, s, t)

    def Fn(R):                      # the law's closed form in FP64, N from the sphere formula
        n = 0.5 + 3.0 * r / (16.0 * R)
        return n ** k / (t - n) if s else n ** k
    Rex = solve_ivp(lambda _t, y: [-Fn(y[0])], (0, T), [1.0], rtol=1e-12, atol=1e-14).y[0, -1]
    ls, g, _ = _gpu_sphere(1.0, r / 8, r + 10 * r / 8)
    ErosionSimulation(ls, law).run(target_fraction=None, max_time=T)
    Rn = _interface_radius(ls, g)
    assert abs(Rn.mean() - Rex) < tol
    assert Rn.max() - Rn.min() < 10 * tol + 1e-4


def test_curvature_sphere_radius(device):
    """Reference tests/test_flows.py:109-116."""
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.flows import GaussCurvatureLaw, MeanCurvatureLaw
    T, h = 0.1, 0.05
    cases = ((MeanCurvatureLaw(1.0), math.sqrt(1 - 2 * T)),
             (GaussCurvatureLaw(1.0), (1 - 3 * T) ** (1 / 3)),
             (GaussCurvatureLaw(0.25, delta=1e-4), (1 - 1.5 * T) ** (2 / 3)))
    for law, Rex in cases:
        ls, g, _ = _gpu_sphere(1.0, h, 12 * h)
        ErosionSimulation(ls, law).run(target_fraction=None, max_time=T)
        Rn = _interface_radius(ls, g)
        assert abs(Rn.mean() - Rex) < 5e-4, law.describe()


def test_history_matches_oracle(device):
    """Same seeded sphere, SVI k=1 s=1 t=0.93 run to a target fraction on GPU and oracle:
    per-step volume fraction within 1e-4, step count and stop reason equal, and the
    final zero sets (marching tetrahedra of both fields) within 0.05 h Hausdorff."""
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.flows import SVIPowerLaw
    r, h = 0.3, 0.05
    ls, g, phi = _gpu_sphere(1.0, h, r + 10 * h)
    sim = ErosionSimulation(ls, SVIPowerLaw(r, 1, 1, 0.93))
    res = sim.run(target_fraction=0.8)
    ref_ls = F.LevelSet(phi.astype(np.float32), g.shape, h)
    ref = F.Simulation(ref_ls, F.SVILaw(r, 1, 1, 0.93))
    reason, steps, _, frac, hist = ref.run(target_fraction=0.8)
    assert res.stop_reason == reason == "target_volume"
    assert abs(sim.initial_volume - ref.initial_volume) < 1e-6 * ref.initial_volume
    assert len(res.history) == len(hist)
    print("history: max fraction difference %.2e over %d steps"
          % (max(abs(a.fraction - b.fraction) for a, b in zip(res.history, hist)), len(hist)), flush=True)
    for a, b in zip(res.history, hist):
        assert abs(a.fraction - b.fraction) < 1e-4
        assert abs(a.dt - b.dt) < 1e-3 * b.dt
        assert a.reinit == b.reinit
    assert abs(res.fraction - 0.8) < 1e-5
    from oracle.surface import marching_tetrahedra
    from svierosion_amd.surface import device_mesh
    v_gpu, f_gpu = device_mesh(ls.flat, ls.band, ls.nb, ls.dims, 0.0)
    v_ref, f_ref = marching_tetrahedra(ref_ls.flat.reshape(g.shape))
    assert len(f_gpu) > 1000 and len(f_ref) > 1000
    dist = sampled_hausdorff(v_gpu, f_gpu, v_ref, f_ref)
    print(f"final surface: sampled Hausdorff distance {dist:.3e} h "
          f"({len(f_gpu)} / {len(f_ref)} triangles)", flush=True)
    assert dist < 0.05


def test_cube_res64_history_matches_oracle(device):
    """The reference's CPU-runnable case (a cube of side 1 as 12 triangles, resolution
    64, SVI k=1 s=1 t=0.93 r=0.1) eroded to 90% volume on the GPU and on the oracle
    from the same signed distance: per-step fraction within 1e-4 and dt within 1e-3
    relative, the same steps, reinitialisations and stop, and final zero sets within
    0.05 h Hausdorff."""
    from oracle.surface import marching_tetrahedra
    from svierosion_amd.api import ErosionParams, make_grid
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.levelset import LevelSet
    from svierosion_amd.sdf import mesh_to_sdf
    from svierosion_amd.shapes import cube
    from svierosion_amd.surface import device_mesh
    mesh = cube(1.0)
    p = ErosionParams(law="svi", radius=0.1, k=1, s=1, t=0.93, resolution=64)
    g = make_grid(mesh, p)
    phi = mesh_to_sdf(mesh, g, band=p.band)
    lo, hi = mesh.bounds
    ls = LevelSet(phi.clone(), g, width=p.band)
    sim = ErosionSimulation(ls, p.make_law(length_scale=float((hi - lo).max())), cfl=p.cfl,
                            feedback_cfl=p.feedback_cfl, reinit_distance=p.reinit_distance,
                            reinit_iterations=p.reinit_iterations)
    res = sim.run(target_fraction=0.9)
    ref_ls = F.LevelSet(phi.cpu().numpy().reshape(g.shape), g.shape, g.spacing, width=p.band)
    ref = F.Simulation(ref_ls, F.SVILaw(0.1, 1, 1, 0.93), cfl=p.cfl, feedback_cfl=p.feedback_cfl,
                       reinit_distance=p.reinit_distance, reinit_iterations=p.reinit_iterations)
    reason, steps, _, frac, hist = ref.run(target_fraction=0.9)
    assert res.stop_reason == reason == "target_volume"
    assert res.steps == steps == len(hist) and any(b.reinit for b in hist)
    err = max(abs(a.fraction - b.fraction) for a, b in zip(res.history, hist))
    for a, b in zip(res.history, hist):
        assert abs(a.dt - b.dt) < 1e-3 * b.dt
        assert a.reinit == b.reinit
    v_gpu, f_gpu = device_mesh(ls.flat, ls.band, ls.nb, ls.dims, 0.0)
    v_ref, f_ref = marching_tetrahedra(ref_ls.flat.reshape(g.shape))
    dist = sampled_hausdorff(v_gpu, f_gpu, v_ref, f_ref)
    print(f"cube res 64: {steps} steps, max fraction difference {err:.2e}, final surfaces "
          f"{dist:.2e} h apart ({len(f_gpu)} / {len(f_ref)} triangles)", flush=True)
    assert err < 1e-4
    assert dist < 0.05


@pytest.mark.parametrize("kind,alpha,delta", [("mean", 1.0, 0.0), ("mean", 0.5, 0.05), ("gauss", 0.25, 1e-3),
                                              ("gauss", 1.0, 0.0)])
def test_curvature_history_matches_oracle(device, kind, alpha, delta):
    """A curvature law (svierosion/flows.py:124-213; velocity svierosion/evolve.py:136-145)
    run to 60% volume on the GPU and on the oracle from the same sphere: per-step
    fraction within 1e-4, dt within 1e-3 relative, the same reinitialisations and
    stop, and final zero sets within 0.05 h.  Covers the regularised alpha < 1 power
    (hardware exp2/log2 on the device) and the parabolic time-step limit."""
    from oracle.surface import marching_tetrahedra
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.flows import GaussCurvatureLaw, MeanCurvatureLaw
    from svierosion_amd.surface import device_mesh
    h = 0.1
    ls, g, phi = _gpu_sphere(1.0, h, 12 * h)
    cls = MeanCurvatureLaw if kind == "mean" else GaussCurvatureLaw
    sim = ErosionSimulation(ls, cls(alpha, 0.0, delta))
    res = sim.run(target_fraction=0.6, max_steps=2000)
    ref_ls = F.LevelSet(phi.astype(np.float32), g.shape, h)
    ref = F.Simulation(ref_ls, F.CurvatureLaw(kind, alpha, delta=delta))
    reason, steps, _, frac, hist = ref.run(target_fraction=0.6, max_steps=2000)
    assert res.stop_reason == reason == "target_volume"
    assert len(res.history) == len(hist) and any(b.reinit for b in hist)
    err = max(abs(a.fraction - b.fraction) for a, b in zip(res.history, hist))
    derr = max(abs(a.dt - b.dt) / b.dt for a, b in zip(res.history, hist))
    for a, b in zip(res.history, hist):
        assert a.reinit == b.reinit
    v_gpu, f_gpu = device_mesh(ls.flat, ls.band, ls.nb, ls.dims, 0.0)
    v_ref, f_ref = marching_tetrahedra(ref_ls.flat.reshape(g.shape))
    dist = sampled_hausdorff(v_gpu, f_gpu, v_ref, f_ref)
    print(f"{kind} alpha={alpha}: {steps} steps, max fraction difference {err:.2e}, max dt difference "
          f"{derr:.2e} (relative), final surfaces {dist:.2e} h apart", flush=True)
    assert err < 1e-4 and derr < 1e-3 and dist < 0.05
    assert abs(res.fraction - 0.6) < 1e-5


@pytest.mark.parametrize("law_name", ["svi", "gauss"])
def test_fused_velocity_rk_is_bit_identical(device, law_name):
    """svimx_velocity_rk (velocity fused with the RK stage) against svimx_velocity +
    svimx_rk_stage: the same run to a target with reinitialisations and the secant
    stop gives identical histories (volume, dt, vmax bit for bit) and identical fields."""
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.flows import GaussCurvatureLaw, SVIPowerLaw
    out = []
    for fused in (True, False):
        if law_name == "svi":
            ls, g, _ = _gpu_sphere(1.0, 0.05, 0.3 + 10 * 0.05)
            sim = ErosionSimulation(ls, SVIPowerLaw(0.3, 2, 1, 0.93))
            target = 0.6              # radius 1 -> 0.843: two reinit distances (1.5 h)
        else:
            ls, g, _ = _gpu_sphere(1.0, 0.1, 1.2)
            sim = ErosionSimulation(ls, GaussCurvatureLaw(0.25, 0.0, 1e-3))
            target = 0.6
        sim.fused = fused
        res = sim.run(target_fraction=target)
        out.append((res, ls.flat.clone(), ls.band.clone()))
    (a, fa, ba), (b, fb, bb) = out
    assert a.stop_reason == b.stop_reason == "target_volume" and any(r.reinit for r in a.history)
    assert [(r.volume, r.dt, r.vmax, r.reinit) for r in a.history] == [(r.volume, r.dt, r.vmax, r.reinit)
                                                                       for r in b.history]
    assert torch.equal(ba, bb) and torch.equal(fa, fb)
    print(f"{law_name}: {len(a.history)} steps identical fused and unfused", flush=True)


@pytest.mark.parametrize("law_name", ["svi", "mean"])
def test_device_dt_matches_host_dt(device, law_name):
    """dt computed on the device (svimx_stable_dt, stages reading it there) against
    the host's stable_dt (svierosion/evolve.py:162-170) on the same fused path: a
    run to a target through reinitialisations and the secant stop, then a run that
    ends on max_time (the clamp of :199-200), give identical histories and fields.
    The device path queues a gated reinitialisation (svimx_reinit_gated) behind the
    steps it predicts to need one; a third run queues it behind every step, so the
    device-side trigger must stay closed on every step the host does not reinitialise."""
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.flows import MeanCurvatureLaw, SVIPowerLaw
    out = []
    for mode in ("device", "host", "gate_every_step"):
        if law_name == "svi":
            ls, g, _ = _gpu_sphere(1.0, 0.05, 0.3 + 10 * 0.05)
            sim = ErosionSimulation(ls, SVIPowerLaw(0.3, 2, 1, 0.93))
        else:
            ls, g, _ = _gpu_sphere(1.0, 0.05, 0.6)
            sim = ErosionSimulation(ls, MeanCurvatureLaw(1.0))
        sim.device_dt = mode != "host"
        if mode == "gate_every_step":
            sim._reinit_likely = lambda: True
        first = sim.run(target_fraction=0.75)
        t_end = sim.time + 2.5 * sim.history[-1].dt   # ends inside the third step
        second = sim.run(target_fraction=None, max_time=t_end)
        out.append((first, second, sim.history, ls.flat.clone(), ls.band.clone()))
    key = lambda h: [(r.step, r.time, r.dt, r.volume, r.vmax, r.vmean, r.qmin, r.qmax, r.capped, r.reinit)
                     for r in h]
    a1, a2, ha, fa, ba = out[0]
    assert a1.stop_reason == "target_volume" and any(r.reinit for r in ha)
    assert a2.stop_reason == "max_time"
    for b1, b2, hb, fb, bb in out[1:]:
        assert b1.stop_reason == a1.stop_reason and b2.stop_reason == a2.stop_reason and a2.steps == b2.steps
        assert str(key(ha)) == str(key(hb))       # str: NaN statistics compare equal
        assert torch.equal(ba, bb) and torch.equal(fa, fb)
    assert ha[-1].dt < ha[-2].dt              # the last step was clamped to max_time
    print(f"{law_name}: {len(ha)} steps ({sum(r.reinit for r in ha)} reinitialisations) identical with dt "
          f"on the device, on the host, and with the gated reinitialisation behind every step", flush=True)


def test_closest_points_newton_steps_match_oracle(device):
    """LevelSet.closest_points with extra Newton steps (svierosion/levelset.py:129-151)
    on a field that is not a distance function (the distorted sphere): the device
    points equal the oracle's to 1e-4 cells for 1, 2 and 3 steps, and the extra steps
    bring the tricubic value at the points closer to zero."""
    from svierosion_amd.grid import Grid
    from svierosion_amd.levelset import LevelSet, tricubic
    from oracle import stencils as S
    phi, origin, h, shape = sphere_phi(1.0, 0.05, 0.6, distort=True)
    phi32 = phi.astype(np.float32)
    ls = LevelSet(torch.from_numpy(phi32).cuda(), Grid(origin, h, shape), width=6.0)
    ref = F.LevelSet(phi32, shape, h, width=6.0)  # the same band and clamped field
    idx = ls.band
    idx_np = idx.cpu().numpy()
    res = []
    for k in (1, 2, 3):
        p = ls.closest_points(idx, newton_steps=k)
        want = S.closest_points(ref.flat, idx_np, shape, h, newton_steps=k)
        err = float(np.abs(p.cpu().numpy() - want).max())
        val = float(tricubic(ls.flat, shape, p).abs().max()) / h
        res.append(val)
        print(f"newton_steps={k}: points {err:.2e} cells from the oracle, max |phi(p)| {val:.2e} h", flush=True)
        assert err < 1e-4
    assert res[1] < 0.5 * res[0] and res[2] < 0.2 * res[1]  # oracle: 4.55 h, 0.85 h, 0.083 h


@pytest.mark.parametrize("law", ["svi", "mean"])
def test_newton_steps_history_matches_oracle(device, law):
    """ErosionSimulation(newton_steps=2) (svierosion/evolve.py:118): closest points
    refined by a tricubic Newton step, their N_r stencils marked for the direct sum,
    the velocity read at them; per-step fraction within 1e-4 and dt within 1e-3 of
    the oracle's run with the same newton_steps, same reinitialisations and stop."""
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.flows import MeanCurvatureLaw, SVIPowerLaw
    h = 0.05 if law == "svi" else 0.1
    r = 0.3
    ls, g, phi = _gpu_sphere(1.0, h, (r + 10 * h) if law == "svi" else 12 * h)
    if law == "svi":
        sim = ErosionSimulation(ls, SVIPowerLaw(r, 1, 1, 0.93), newton_steps=2)
        ref_law, target = F.SVILaw(r, 1, 1, 0.93), 0.85
    else:
        sim = ErosionSimulation(ls, MeanCurvatureLaw(1.0, 0.0, 0.0), newton_steps=2)
        ref_law, target = F.CurvatureLaw("mean", 1.0, delta=0.0), 0.7
    res = sim.run(target_fraction=target, max_steps=2000)
    ref = F.Simulation(F.LevelSet(phi.astype(np.float32), g.shape, h), ref_law, newton_steps=2)
    reason, steps, _, frac, hist = ref.run(target_fraction=target, max_steps=2000)
    assert res.stop_reason == reason == "target_volume"
    assert len(res.history) == len(hist)
    err = max(abs(a.fraction - b.fraction) for a, b in zip(res.history, hist))
    derr = max(abs(a.dt - b.dt) / b.dt for a, b in zip(res.history, hist))
    print(f"{law} newton_steps=2: {steps} steps, max fraction difference {err:.2e}, dt {derr:.2e} relative",
          flush=True)
    for a, b in zip(res.history, hist):
        assert a.reinit == b.reinit
    assert err < 1e-4 and derr < 1e-3
