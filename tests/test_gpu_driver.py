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
