"""Every benchmark workload run on the GPU and checked:

* fragment, res 256 (icosphere(6) + 16 SH terms, SVI k=1 s=1 t=0.93, r=2): the
  first two steps against the oracle (same SDF field), and erode() to 90% with
  the property checks of the reference's driver test (tests/test_flows.py:120-129);
* scan, res 400 (1,310,720 triangles, SVI k=2 s=1 t=0.95, r=2): SDF at sampled band
  nodes against an FP64 brute force over nearby triangles, N_r at sampled nodes
  against the FP64 direct sum, and erode() with its exact stops at 0.97, 0.94, 0.9;
* the batched fragment mix and the served Poisson requests (first six jobs of
  each): continuous batching equals sequential erode() bit for bit.
(The cube case, res 64, is covered by tests/test_gpu_flow.py and
tests/test_gpu_ops.py.)
"""

import math

import numpy as np
import pytest
import torch

from oracle import flow as F
from oracle import stencils as S

pytestmark = pytest.mark.gpu


def _mesh_volume_ratio(res):
    return res.mesh.volume() / res.initial_mesh.volume()


def _check_erode_result(res, targets, tol=1e-6):
    """Exact stops (to the regula-falsi tolerance), watertight outputs, mesh volume
    within 2e-3 of the level set's fraction, finite vertex fields."""
    assert res.stop_reason == "target_volume"
    fracs = [f for f, _, _ in res.snapshots] + [res.fraction]
    assert len(fracs) == len(targets)
    for f, t in zip(fracs, targets):
        assert abs(f - t) <= tol * 1.01, (f, t)
    for _, _, m in res.snapshots:
        assert m.is_watertight()
    assert res.mesh.is_watertight() and res.mesh.is_consistently_oriented()
    ratio = _mesh_volume_ratio(res)
    assert abs(ratio - res.fraction) < 2e-3, (ratio, res.fraction)
    for k, v in res.mesh.vertex_data.items():
        assert np.isfinite(v).all(), k
    return fracs, ratio


# ------------------------------------------------------------------ fragment, res 256
def _fragment():
    from svierosion_amd.api import ErosionParams
    from svierosion_amd.shapes import sh_fragment
    mesh = sh_fragment(level=6, radius=50.0, seed=0)
    return mesh, ErosionParams(law="svi", radius=2.0, k=1, s=1, t=0.93, resolution=256, target_fraction=0.9)


def test_fragment_res256_first_steps_match_oracle(device):
    """Two steps of the 81,920-triangle fragment at res 256 on the GPU and on the oracle
    from the same signed distance: initial volume within 1e-6 relative, per-step
    fraction within 1e-4 and dt within 1e-3 relative."""
    from svierosion_amd.api import make_grid
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.levelset import LevelSet
    from svierosion_amd.sdf import mesh_to_sdf
    mesh, p = _fragment()
    assert len(mesh.faces) == 81920
    g = make_grid(mesh, p)
    phi = mesh_to_sdf(mesh, g, band=p.band)
    lo, hi = mesh.bounds
    law = p.make_law(length_scale=float((hi - lo).max()))
    sim = ErosionSimulation(LevelSet(phi.clone(), g, width=p.band), law)
    res = sim.run(target_fraction=None, max_steps=2)
    ref = F.Simulation(F.LevelSet(phi.cpu().numpy().reshape(g.shape), g.shape, g.spacing, width=p.band),
                       F.SVILaw(2.0, 1, 1, 0.93))
    _, steps, _, _, hist = ref.run(target_fraction=None, max_steps=2)
    assert steps == 2 == len(res.history)
    ev = abs(sim.initial_volume - ref.initial_volume) / ref.initial_volume
    ef = max(abs(a.fraction - b.fraction) for a, b in zip(res.history, hist))
    ed = max(abs(a.dt - b.dt) / b.dt for a, b in zip(res.history, hist))
    print(f"fragment res 256 ({g.shape}): initial volume {ev:.2e} relative, fraction {ef:.2e}, dt {ed:.2e} "
          f"relative over 2 steps", flush=True)
    assert ev < 1e-6 and ef < 1e-4 and ed < 1e-3


def test_fragment_res256_erode_to_target(device):
    from svierosion_amd.api import erode
    mesh, p = _fragment()
    res = erode(mesh, p)
    fracs, ratio = _check_erode_result(res, [0.9])
    print(f"fragment res 256: {len(res.history)} steps, stop {fracs[-1]:.8f}, mesh volume ratio {ratio:.5f}",
          flush=True)


# ------------------------------------------------------------------ scan, res 400
@pytest.fixture(scope="module")
def scan():
    from svierosion_amd.api import ErosionParams, make_grid
    from svierosion_amd.shapes import noisy_ellipsoid
    mesh = noisy_ellipsoid(level=8, axes=(60.0, 45.0, 30.0), amplitudes=(3.0, 1.5, 0.75, 0.375),
                           base_frequency=2.0, seed=0)
    p = ErosionParams(law="svi", radius=2.0, k=2, s=1, t=0.95, resolution=400, target_fraction=0.9,
                      snapshots=(0.97, 0.94))
    return mesh, p, make_grid(mesh, p)


def test_scan_res400_sdf_sampled_against_brute_force(device, scan):
    """20k sampled band nodes of the 1.3M-triangle scan: within the exact near field
    (distance <= 2 cells, svierosion/sdf.py:100-145) the distance equals the FP64
    brute force over nearby triangles within 1e-6 h; beyond it the distance comes
    from ring propagation and star refinement (svierosion/sdf.py:148-265), which take
    the closest of the triangles their neighbours found: a distance to an actual
    triangle, so never below the true distance, and equal to it unless the true
    closest triangle was missed.  Measured: 99.99% of 13k propagated nodes within
    1e-6 h, one 2.3e-3 h above (the same algorithm as the FP64 oracle, which the
    fragment SDF test matches to 1e-6 h)."""
    from scipy.spatial import cKDTree
    from svierosion_amd.sdf import mesh_to_sdf
    mesh, p, g = scan
    assert len(mesh.faces) == 1310720
    phi, dist, _ = mesh_to_sdf(mesh, g, band=p.band, return_aux=True)
    d = dist.reshape(-1)
    band = torch.nonzero(d < 5.99).squeeze(1)
    gen = torch.Generator().manual_seed(1)
    sel = band[torch.randperm(band.numel(), generator=gen)[:20000].to(band.device)]
    dg = d[sel].double().cpu().numpy()
    pts = g.node_coords(sel).cpu().numpy()
    tri = mesh.triangles()
    cen = tri.mean(1)
    rmax = float(np.linalg.norm(tri - cen[:, None, :], axis=2).max())
    tree = cKDTree(cen)
    exact = np.empty(len(pts))
    for q, cand in enumerate(tree.query_ball_point(pts, dg * g.spacing + rmax + 1e-9)):
        t = tri[cand]
        exact[q] = math.sqrt(S.point_triangle_distance2(pts[q][None], t[:, 0], t[:, 1], t[:, 2]).min())
    exact /= g.spacing
    near = dg <= 2.0
    err_near = np.abs(dg - exact)[near].max()
    over = (dg - exact)[~near]
    print(f"scan SDF: {near.sum()} near-field nodes within {err_near:.2e} h of the brute force; "
          f"{(~near).sum()} propagated nodes: {np.mean(np.abs(over) < 1e-6):.4f} within 1e-6 h, "
          f"max excess {over.max():.2e} h", flush=True)
    assert err_near < 1e-6
    assert over.min() > -1e-6
    assert np.mean(np.abs(over) < 1e-6) > 0.999 and over.max() < 0.01


def test_scan_res400_nr_sampled_against_fp64_direct_sum(device, scan):
    """N_r of the initial scan at 2,000 sampled interface nodes (the direct summation the
    steps use) against the FP64 direct sum of the same occupancy, within 2e-5."""
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.levelset import LevelSet
    from svierosion_amd.sdf import mesh_to_sdf
    mesh, p, g = scan
    lo, hi = mesh.bounds
    law = p.make_law(length_scale=float((hi - lo).max()))
    ls = LevelSet(mesh_to_sdf(mesh, g, band=p.band), g, width=p.band)
    sim = ErosionSimulation(ls, law)
    iface = ls.interface_nodes()
    gen = torch.Generator().manual_seed(2)
    nodes = iface[torch.randperm(iface.numel(), generator=gen)[:2000].to(iface.device)]
    frac = ls.dense_fractions()
    got = sim.conv.exterior_fraction(frac, nodes).reshape(-1)[nodes.long()].double().cpu().numpy()
    ijk = np.stack(np.unravel_index(nodes.long().cpu().numpy(), g.shape), 1)
    w = F.ball_cell_weights(law.radius, g.spacing)
    ref = F.exterior_fraction_direct(frac.cpu().numpy().reshape(g.shape), w, g.spacing, law.radius, ijk)
    err = np.abs(got - ref).max()
    print(f"scan N_r at {len(ijk)} interface nodes: max |GPU - FP64 direct sum| {err:.2e}", flush=True)
    assert err < 2e-5


def test_scan_res400_erode_exact_stops(device, scan):
    from svierosion_amd.api import erode
    mesh, p, _ = scan
    res = erode(mesh, p)
    fracs, ratio = _check_erode_result(res, [0.97, 0.94, 0.9])
    print(f"scan res 400: {len(res.history)} steps, stops {[round(f, 8) for f in fracs]}, final mesh volume "
          f"ratio {ratio:.5f}", flush=True)


# ------------------------------------------------------------------ batched and served workloads
def test_fragment_mix_prefix_batched_equals_sequential(device):
    """The first six jobs of the 64-fragment mix (SVI, mean- and Gauss-curvature jobs at
    res 123-229, one of them ~28k steps) batched in one engine equal sequential erode()."""
    from svierosion_amd.batch import Request, serve, serve_sequential
    from svierosion_amd.workloads import fragment_mix
    from tests.test_gpu_serve import _check_same
    reqs = [Request(m, p) for m, p in fragment_mix(6)]
    assert {r.params.law for r in reqs} == {"svi", "mean", "gauss"}
    mix = serve(reqs, workers=2, schedule="lockstep")  # the batch schedule of erode_batch
    seq = serve_sequential(reqs)
    assert len(mix.done) == len(seq.done) == 6
    _check_same(mix, seq)
    print(f"fragment mix prefix: batched {mix.summary()} / sequential {seq.summary()}", flush=True)


def test_poisson_requests_prefix_served_equals_sequential(device):
    """The first six of the 512 Poisson requests (res 68-311, targets in [0.85, 0.97])
    arriving at 2/s, continuously batched, equal sequential erode()."""
    from svierosion_amd.batch import serve, serve_sequential
    from svierosion_amd.workloads import poisson_requests
    from tests.test_gpu_serve import _check_same
    reqs = poisson_requests(6, rate=2.0)
    mix = serve(reqs, workers=2)
    seq = serve_sequential(reqs)
    assert len(mix.done) == len(seq.done) == 6
    _check_same(mix, seq)
    assert all(c.admitted >= c.arrival for c in mix.completions)
    print(f"served prefix: {mix.summary()} / sequential {seq.summary()}", flush=True)
