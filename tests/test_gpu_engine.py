"""Batch engine: every job's history and field are bit-identical to the
single-job path (ErosionSimulation.run, svierosion/evolve.py:182-250), alone
and in a mix of jobs with other grids and laws, with and without HIP graphs."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(name):
    from svierosion_amd.api import ErosionParams
    from svierosion_amd.shapes import cube, icosphere
    name = name.removesuffix("_lds").removesuffix("_y2").removesuffix("_y4")
    if name == "svi_cube":
        return cube(1.0), ErosionParams(law="svi", radius=0.2, k=1, s=1, t=0.9, resolution=40), 0.93, 400
    if name == "svi_sphere_k2":
        return icosphere(3, 0.8), ErosionParams(law="svi", radius=0.25, k=2, s=0, resolution=36), 0.95, 400
    if name == "mean":
        return icosphere(3, 0.8), ErosionParams(law="mean", alpha=1.0, resolution=32), 0.96, 400
    if name == "gauss":
        return cube(1.0), ErosionParams(law="gauss", alpha=0.25, resolution=24), 0.97, 40
    if name == "svi_fft":
        return cube(1.0), ErosionParams(law="svi", radius=0.2, k=1, s=0, resolution=40), 0.95, 400
    raise KeyError(name)


_PHI = {}


def _sim(name):
    """A fresh simulation of case `name` (the signed distance is computed once and cloned)."""
    from svierosion_amd.api import make_grid
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.levelset import LevelSet
    from svierosion_amd.sdf import mesh_to_sdf
    mesh, p, target, max_steps = _case(name)
    base = name.removesuffix("_lds").removesuffix("_y2").removesuffix("_y4")
    grid = make_grid(mesh, p)
    if base not in _PHI:
        _PHI[base] = mesh_to_sdf(mesh, grid, band=p.band)
    lo, hi = mesh.bounds
    law = p.make_law(length_scale=float((hi - lo).max()))
    ls = LevelSet(_PHI[base].clone(), grid, width=p.band)
    sim = ErosionSimulation(ls, law, cfl=p.cfl, feedback_cfl=p.feedback_cfl, reinit_distance=p.reinit_distance,
                            reinit_iterations=p.reinit_iterations, nr_method="fft" if name == "svi_fft" else "auto")
    if sim.conv is not None and sim.conv.method == "direct":
        sim.conv.staging, sim.conv.lines = "l1", 1
        if name.endswith("_lds"):
            sim.conv.staging = "lds"
        elif name.endswith("_y2") or name.endswith("_y4"):
            sim.conv.lines = int(name[-1])
    return sim, target, max_steps


def _same(a, b):
    return (math.isnan(a) and math.isnan(b)) or a == b


def _assert_same_run(sa, sb):
    import torch
    assert len(sa.history) == len(sb.history) > 0
    for ra, rb in zip(sa.history, sb.history):
        for f in ("step", "time", "dt", "volume", "fraction", "vmax", "vmean", "qmin", "qmax", "capped", "reinit"):
            assert _same(getattr(ra, f), getattr(rb, f)), (ra.step, f, getattr(ra, f), getattr(rb, f))
    la, lb = sa.ls, sb.ls
    assert la.nb == lb.nb and la.inside_outside == lb.inside_outside
    assert torch.equal(la.band, lb.band)
    assert torch.equal(la.flat, lb.flat)
    assert sa.time == sb.time and sa.steps == sb.steps


def _engine_run(names, graphs, group_nodes=None, schedule=None):
    from svierosion_amd.engine import BatchEngine
    eng = BatchEngine(max_jobs=8, graphs=graphs)
    if group_nodes is not None:
        eng.group_nodes(group_nodes)
    if schedule is not None:
        eng.schedule(*schedule)
    jobs = {}
    for nm in names:
        sim, target, max_steps = _sim(nm)
        j = eng.add(sim)
        eng.run(j, target_fraction=target, max_steps=max_steps)
        jobs[nm] = j
    eng.run_all()
    out = {}
    for nm, j in jobs.items():
        sim = eng.sync(j)
        out[nm] = (sim, eng.stop_reason(j))
    eng.close()
    return out


@pytest.mark.parametrize("name", ["svi_cube", "svi_sphere_k2", "mean", "gauss", "svi_fft", "svi_cube_lds",
                                  "svi_cube_y2", "svi_sphere_k2_y4"])
def test_engine_job_equals_single_job_path(device, name):
    sim, target, max_steps = _sim(name)
    r = sim.run(target_fraction=target, max_steps=max_steps)
    got = _engine_run([name], graphs=False)[name]
    assert got[1] == r.stop_reason
    _assert_same_run(sim, got[0])
    if name != "gauss":
        assert r.stop_reason == "target_volume"
        assert abs(sim.history[-1].fraction - target) < 5e-6  # the regula falsi ran
    if name == "svi_cube":
        assert any(rec.reinit for rec in sim.history)


@pytest.mark.parametrize("shape", ["lds", "y2", "y4"])
def test_nr_shapes_are_result_neutral_in_the_engine(device, shape):
    """The engine's multi-line N_r families reproduce the single-job path with
    one-line windows read through the L1, bit for bit; its LDS-staged family
    (which sums in its own order) reproduces the single-job LDS kernel."""
    sim, target, max_steps = _sim("svi_sphere_k2")
    assert sim.conv.staging == "l1" and sim.conv.lines == 1
    if shape == "lds":
        sim.conv.staging = "lds"
    r = sim.run(target_fraction=target, max_steps=max_steps)
    name = "svi_sphere_k2_" + shape
    got = _engine_run([name], graphs=True)[name]
    assert got[0].conv.shape_name.endswith({"lds": "/lds", "y2": "x2", "y4": "x4"}[shape])
    assert got[1] == r.stop_reason
    _assert_same_run(sim, got[0])


def test_mix_equals_alone_with_graphs(device):
    names = ["svi_cube", "mean", "svi_sphere_k2_lds", "gauss", "svi_fft"]
    mix = _engine_run(names, graphs=True)
    for nm in names:
        alone = _engine_run([nm], graphs=False)[nm]
        assert mix[nm][1] == alone[1]
        _assert_same_run(mix[nm][0], alone[0])


@pytest.mark.parametrize("group_nodes", [1, 40000])
def test_launch_groups_are_result_neutral(device, group_nodes):
    """A tick split into launch groups (every job alone, or groups of at most 40k
    band nodes; svimx_engine_group_nodes) gives the same histories and fields as
    one group, with graphs (captured per set and group shape) and through the
    secant retries and reinitialisations."""
    names = ["svi_cube", "mean", "svi_sphere_k2_y2", "gauss", "svi_fft"]
    one = _engine_run(names, graphs=True, group_nodes=0)
    split = _engine_run(names, graphs=True, group_nodes=group_nodes)
    for nm in names:
        assert split[nm][1] == one[nm][1]
        _assert_same_run(split[nm][0], one[nm][0])
    sizes = {nm: one[nm][0].ls.nb for nm in names}
    print(f"groups of <= {group_nodes} band nodes: identical to one group; band sizes {sizes}", flush=True)


def test_short_schedule_serialises_long_jobs(device):
    """SVIMX_SCHED_SHORT with a zero work bound: behind the oldest job only jobs
    fewer than 4 steps into their run join a tick, so after their first steps the
    jobs run one after another in the order they were added; every history and
    field equals the lockstep run."""
    names = ["svi_cube", "mean", "svi_sphere_k2_y2", "gauss"]
    one = _engine_run(names, graphs=True, schedule=("lockstep",))
    short = _engine_run(names, graphs=True, schedule=("short", 10**9, 0.0, 0.0))
    for nm in names:
        assert short[nm][1] == one[nm][1]
        _assert_same_run(short[nm][0], one[nm][0])


def test_engine_rerun_with_lower_target_continues(device):
    """Snapshots: a job stopped at one target continues to the next one."""
    from svierosion_amd.engine import BatchEngine
    sim_a, _, _ = _sim("svi_cube")
    sim_a.run(target_fraction=0.97)
    sim_a.run(target_fraction=0.94)
    sim_b, _, _ = _sim("svi_cube")
    eng = BatchEngine(max_jobs=2)
    j = eng.add(sim_b)
    for tf in (0.97, 0.94):
        eng.run(j, target_fraction=tf)
        eng.run_all()
        assert eng.stop_reason(j) == "target_volume"
    eng.sync(j)
    eng.close()
    _assert_same_run(sim_a, sim_b)
    assert abs(sim_b.history[-1].fraction - 0.94) < 5e-6
