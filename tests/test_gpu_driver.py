"""erode() end to end on the GPU (reference tests/test_flows.py:118-134 semantics)."""

import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_stops_exactly_at_target_volume(device):
    from svierosion_amd import ErosionParams, erode
    from svierosion_amd.shapes import cube
    m = cube(1.0)
    res = erode(m, ErosionParams(law="svi", radius=0.2, k=1, s=1, t=0.9, resolution=60, target_fraction=0.93))
    assert res.stop_reason == "target_volume"
    assert abs(res.fraction - 0.93) < 5e-6
    assert abs(res.mesh.volume() / m.volume() - 0.93) < 2e-3
    assert res.mesh.is_watertight()
    for key in ("depth", "N", "speed"):
        assert key in res.mesh.vertex_data
    assert np.all(res.mesh.vertex_data["depth"] >= 0)
    assert res.history and res.history[-1].fraction == res.fraction
    assert set(res.timings) == {"sdf", "evolve", "mesh"}


def test_rejects_t_below_max_n(device):
    from svierosion_amd import ErosionParams, erode
    from svierosion_amd.shapes import cube
    with pytest.raises(ValueError):
        erode(cube(1.0), ErosionParams(law="svi", radius=0.2, k=1, s=1, t=0.8, resolution=40))


def test_snapshots_and_ply_output(device):
    from svierosion_amd import ErosionParams, erode, read_ply, write_ply
    from svierosion_amd.shapes import cube
    res = erode(cube(1.0), ErosionParams(law="svi", radius=0.2, k=1, s=0, resolution=40, target_fraction=0.94,
                                         snapshots=(0.97,)))
    assert res.stop_reason == "target_volume"
    assert len(res.snapshots) == 1 and abs(res.snapshots[0][0] - 0.97) < 1e-5
    assert res.snapshots[0][2].is_watertight()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "out.ply")
        write_ply(p, res.mesh)
        back = read_ply(p)
        assert back.faces.shape == res.mesh.faces.shape
        assert "N" in back.vertex_data


def test_curvature_law_erode_annotations(device):
    from svierosion_amd import ErosionParams, erode
    from svierosion_amd.shapes import icosphere
    res = erode(icosphere(3, 0.8), ErosionParams(law="mean", alpha=1.0, resolution=32, target_fraction=0.95))
    assert res.stop_reason == "target_volume"
    for key in ("depth", "H", "K", "speed"):
        assert key in res.mesh.vertex_data


def test_stop_reasons_max_steps_and_max_time(device):
    import torch
    from svierosion_amd.api import ErosionParams, make_grid
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.flows import SVIPowerLaw
    from svierosion_amd.levelset import LevelSet
    from svierosion_amd.sdf import mesh_to_sdf
    from svierosion_amd.shapes import cube
    m = cube(1.0)
    p = ErosionParams(law="svi", radius=0.2, resolution=40)
    g = make_grid(m, p)
    ls = LevelSet(mesh_to_sdf(m, g), g)
    sim = ErosionSimulation(ls, SVIPowerLaw(0.2, 1, 0))
    r = sim.run(target_fraction=0.5, max_steps=1)
    assert r.stop_reason == "max_steps" and r.steps == 1
    r = sim.run(target_fraction=None, max_time=sim.time + 1e-3)
    assert r.stop_reason == "max_time"
    assert abs(sim.time - sim.history[-1].time) < 1e-15


@pytest.mark.parametrize("law_kw", [dict(law="svi", radius=0.2, k=2, s=1, t=0.95),
                                    dict(law="mean", alpha=1.0), dict(law="gauss", alpha=0.25)])
def test_vertex_fields_follow_the_reference_definition(device, law_kw):
    """N / H, K and speed at the output vertices (one fused pass, svimx_annotate_points)
    against the definition of svierosion/api.py:171-194 rebuilt from separate pieces:
    tricubic interpolation of the full dense N_r (SVI) or of dense k1, k2 fields that
    are zero off the band (curvature), and the oracle's law in FP64."""
    import torch
    from oracle import flow as F
    from svierosion_amd import ErosionParams, erode
    from svierosion_amd.levelset import tricubic
    from svierosion_amd.shapes import icosphere
    res = erode(icosphere(4, 0.8), ErosionParams(resolution=40, target_fraction=0.96, **law_kw), keep_state=True)
    sim, ls, mesh = res.simulation, res.level_set, res.mesh
    p = torch.as_tensor(ls.grid.to_index(mesh.vertices), dtype=torch.float32, device=ls.device)
    law = sim.law
    vd = mesh.vertex_data
    if law.requires_svi:
        N = tricubic(sim.svi_field().reshape(-1), ls.dims, p).double().cpu().numpy()
        ref = F.SVILaw(law.radius, law.k, law.s, law.t).speed(N)
        err_q = np.abs(vd["N"] - N).max()
        assert err_q < 2e-6
    else:
        k1, k2 = ls.principal_curvatures(ls.band)
        dense = torch.zeros(ls.grid.size, dtype=torch.float32, device=ls.device)
        dense[ls.band.long()] = k1
        K1 = tricubic(dense, ls.dims, p).double().cpu().numpy()
        dense.zero_()
        dense[ls.band.long()] = k2
        K2 = tricubic(dense, ls.dims, p).double().cpu().numpy()
        kind = "mean" if law.name == "mean" else "gauss"
        ref, _ = F.CurvatureLaw(kind, law.alpha, law.eps, law.delta).speed_and_diffusivity(K1, K2)
        scale = max(np.abs(K1).max(), np.abs(K2).max())
        err_q = max(np.abs(vd["H"] - 0.5 * (K1 + K2)).max() / scale, np.abs(vd["K"] - K1 * K2).max() / scale ** 2)
        assert err_q < 1e-5
    err_v = np.abs(vd["speed"] - ref).max() / max(np.abs(ref).max(), 1e-30)
    print(f"{law_kw['law']}: {len(mesh.vertices)} vertices, field error {err_q:.2e}, speed error {err_v:.2e}")
    assert err_v < 1e-5


def test_small_grid_job_in_the_engine_equals_erode(device):
    """A job whose band is a large share of its grid (cube res 64: 262k band nodes of
    90^3) in the batch engine equals erode(): the engine grows its reinitialisation
    work buffers by a quarter over the request, past the node count, and the
    reinitialisation uses them up to the grid size (it refused the compaction
    before)."""
    from svierosion_amd import ErosionParams, erode
    from svierosion_amd.batch import Request, serve
    from svierosion_amd.shapes import cube
    p = ErosionParams(law="svi", radius=0.1, k=1, s=1, t=0.93, resolution=64, target_fraction=0.9,
                      snapshots=(0.95,))
    rep = serve([Request(cube(1.0), p)], solo_tail=False)  # every step in the engine
    assert rep.completions[0].error is None, rep.completions[0].error
    eng = rep.completions[0].result
    run = erode(cube(1.0), p)
    assert eng.stop_reason == run.stop_reason == "target_volume"
    assert [(r.volume, r.dt, r.reinit) for r in eng.history] == [(r.volume, r.dt, r.reinit) for r in run.history]
    assert any(r.reinit for r in eng.history)
    assert [f for f, _, _ in eng.snapshots] == [f for f, _, _ in run.snapshots]
    assert np.array_equal(eng.mesh.vertices, run.mesh.vertices) and np.array_equal(eng.mesh.faces, run.mesh.faces)
