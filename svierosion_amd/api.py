"""High-level interface: erode a triangle mesh (API of svierosion/api.py).

``erode(mesh, ErosionParams)`` keeps the reference's signature, parameters,
result fields, stop reasons and error types.  The pipeline is the same
(svierosion/api.py:197-253): grid sizing, mesh SDF, level set, simulation,
t > max N_r check, evolution to each snapshot target, surface extraction and
per-vertex annotation -- with every numerical stage on the GPU through
libsvimx.
"""

from __future__ import annotations

import ctypes as C
import time
import warnings
from dataclasses import asdict, dataclass, field

import numpy as np
import torch

from . import _abi
from .evolve import ErosionSimulation, StepRecord
from .flows import GaussCurvatureLaw, MeanCurvatureLaw, SpeedLaw, SVIPowerLaw
from .grid import Grid, get_device
from .levelset import MARGIN, LevelSet, tricubic
from .meshio import TriMesh
from .sdf import mesh_to_sdf, triangle_distances, triangle_frames
from .surface import extract_surface


@dataclass
class ErosionParams:
    """All parameters of an erosion run (lengths in mesh units); ``spacing`` beats ``resolution``."""

    law: str = "svi"
    radius: float | None = None
    k: int = 1
    s: int = 0
    t: float | None = None
    eps: float = 0.0
    max_speed_ratio: float = 100.0
    alpha: float = 1.0
    delta: float | None = None
    target_fraction: float = 0.9
    snapshots: tuple = ()
    spacing: float | None = None
    resolution: int = 200
    band: float = 6.0
    cfl: float = 0.5
    feedback_cfl: float = 1.0
    parabolic_cfl: float = 0.25
    reinit_distance: float = 1.5
    reinit_iterations: int = 20
    project_vertices: bool = True
    device: str | None = None
    max_steps: int = 1_000_000

    def make_law(self, length_scale: float = 1.0) -> SpeedLaw:
        """Speed law; ``length_scale`` sets the automatic delta (svierosion/api.py:54-66)."""
        if self.law == "svi":
            if self.radius is None:
                raise ValueError("the SVI law needs a radius")
            return SVIPowerLaw(self.radius, self.k, self.s, self.t, self.max_speed_ratio, self.eps)
        if self.law in ("mean", "mean_curvature"):
            d = self.delta if self.delta is not None else (1.0 / length_scale if self.alpha < 1 else 0.0)
            return MeanCurvatureLaw(self.alpha, self.eps, d)
        if self.law in ("gauss", "gauss_curvature"):
            d = self.delta if self.delta is not None else (length_scale ** -2 if self.alpha < 1 else 0.0)
            return GaussCurvatureLaw(self.alpha, self.eps, d)
        raise ValueError(f"unknown law {self.law!r}")


@dataclass
class ErosionResult:
    mesh: TriMesh
    snapshots: list = field(default_factory=list)
    history: list = field(default_factory=list)
    grid: Grid | None = None
    stop_reason: str = ""
    time: float = 0.0
    fraction: float = 1.0
    initial_volume_grid: float = 0.0
    initial_volume_mesh: float = 0.0
    timings: dict = field(default_factory=dict)
    params: dict = field(default_factory=dict)
    initial_mesh: TriMesh | None = None
    level_set: LevelSet | None = None
    simulation: ErosionSimulation | None = None

    def summary(self) -> str:
        """One line: outcome, grid and phase times."""
        t = self.timings
        phases = ", ".join(f"{k} {t.get(k, 0.0):.2f} s" for k in ("sdf", "evolve", "mesh"))
        return (f"{self.stop_reason} after {len(self.history)} steps (t = {self.time:.6g}), "
                f"{100.0 * self.fraction:.3f}% of the volume left; {self.grid.describe()}; {phases}")


def make_grid(mesh: TriMesh, params: ErosionParams) -> Grid:
    """h = L/resolution (or spacing); padding max((band+MARGIN+2)h, r+3h) (svierosion/api.py:94-101)."""
    lo, hi = mesh.bounds
    ext = float((hi - lo).max())
    h = params.spacing if params.spacing else ext / params.resolution
    pad = (params.band + MARGIN + 2) * h
    if params.law == "svi":
        pad = max(pad, params.radius + 3 * h)
    return Grid.around(lo, hi, h, pad)


def point_mesh_distance(points: np.ndarray, mesh: TriMesh, max_edge: float | None = None, k: int = 16,
                        device=None, exact: bool = True, grid: Grid | None = None,
                        triangles: dict | None = None) -> np.ndarray:
    """Unsigned distance from points to a mesh (svierosion/api.py:104-168).

    Exact on the device (``svimx_point_mesh_distance``): the mesh's triangles,
    split to edges <= 8 cells of ``grid`` and stored in triangle-local frames,
    are binned into 4-cell buckets and every point searches rings of buckets
    until a lower bound proves the minimum.  ``grid`` defaults to a frame
    around the mesh and the points with the mesh's median edge as the cell;
    ``triangles`` reuses frames already on the device.  ``max_edge``, ``k`` and
    ``exact`` of the reference's k-d tree search are accepted and have no
    effect (the result is always the certified minimum)."""
    dev = _abi.require_device(get_device(device))
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    if grid is None:
        e = np.linalg.norm(mesh.vertices[mesh.faces[:, 1]] - mesh.vertices[mesh.faces[:, 0]], axis=1)
        lo = np.minimum(mesh.vertices.min(0), points.min(0) if len(points) else mesh.vertices.min(0))
        hi = np.maximum(mesh.vertices.max(0), points.max(0) if len(points) else mesh.vertices.max(0))
        cell = max(float(np.median(e)) if len(e) else 1.0, float((hi - lo).max()) / 512.0, 1e-12)
        shape = tuple(int(v) for v in np.ceil((hi - lo) / cell) + 3)
        grid = Grid(tuple(lo - cell), cell, shape)
        triangles = None
    if triangles is None:
        triangles = triangle_frames(mesh, grid, dev)
    pidx = torch.as_tensor((points - np.asarray(grid.origin)) / grid.spacing, dtype=torch.float32, device=dev)
    d = triangle_distances(triangles, grid, pidx)
    return d.double().cpu().numpy() * grid.spacing


def _annotate(sim: ErosionSimulation, mesh: TriMesh, initial: TriMesh, law: SpeedLaw,
              triangles: dict | None = None) -> None:
    """Per-vertex fields of an output surface: ``depth`` below the input mesh, then
    ``N`` (SVI laws) or ``H`` and ``K`` (curvature laws) and the law's ``speed``
    there (the field set of svierosion/api.py:171-194).

    One device pass (``svimx_annotate_points``) interpolates N_r -- evaluated at
    the stencils of the vertices only -- or the dense (k1, k2) pairs of the final
    field tricubically and evaluates the law with the velocity pass's functions."""
    if len(mesh.vertices) == 0:
        return
    ls = sim.ls
    grid = ls.grid
    mesh.vertex_data["depth"] = point_mesh_distance(mesh.vertices, initial, device=ls.device, grid=grid,
                                                    triangles=triangles).astype(np.float32)
    p = torch.as_tensor(grid.to_index(mesh.vertices), dtype=torch.float32, device=ls.device).contiguous()
    n = int(p.shape[0])
    q1, q2, speed = (torch.empty(n, dtype=torch.float32, device=ls.device) for _ in range(3))
    st = _abi.stream()
    field = k12 = None
    if law.requires_svi:
        field = sim.svi_field_at(p)
    else:
        # the pairs array is zero off the band; refresh it on the final field
        k12 = sim.k12
        _abi.check(_abi.lib.svimx_curvatures(st, _abi.ptr(ls.flat), _abi.ptr(ls._band), ls.nb, *ls.dims, ls.h,
                                             _abi.ptr(k12)), "curvatures")
    rec = law.to_abi()
    _abi.check(_abi.lib.svimx_annotate_points(st, C.byref(rec), _abi.ptr(field), _abi.ptr(k12), *ls.dims,
                                              _abi.ptr(p), n, _abi.ptr(q1), _abi.ptr(q2), _abi.ptr(speed)),
               "vertex fields")
    host = [t.cpu().numpy() for t in (q1, q2, speed)]
    if law.requires_svi:
        mesh.vertex_data["N"] = host[0]
    else:
        mesh.vertex_data["H"], mesh.vertex_data["K"] = host[0], host[1]
    mesh.vertex_data["speed"] = host[2]


def initial_max_nr(sim: ErosionSimulation) -> float:
    """max N_r at the closest points of the initial interface nodes (svierosion/api.py:220-223)."""
    ls = sim.ls
    cp = ls.closest_points(ls.interface_nodes())
    if cp.shape[0] == 0:
        return 0.0
    N = sim.svi_field_at(cp)
    return float(tricubic(N.reshape(-1), ls.dims, cp).max())


@dataclass
class ErosionJob:
    """One erosion between set-up and result: the pieces of erode() shared with
    the batched and served drivers (svierosion_amd.batch)."""

    mesh: TriMesh
    params: ErosionParams
    law: SpeedLaw
    grid: Grid
    ls: LevelSet
    sim: ErosionSimulation
    targets: list
    timings: dict = field(default_factory=dict)
    snaps: list = field(default_factory=list)
    next_target: int = 0
    stop_reason: str = ""
    fraction: float = 1.0
    tmesh: float = 0.0
    triangles: dict = field(default_factory=dict)  # device split triangles of the input mesh


def prepare_job(mesh: TriMesh, params: ErosionParams, verbose: bool = False) -> ErosionJob:
    """Grid, mesh SDF, level set, simulation and the t > max N_r check (svierosion/api.py:197-228)."""
    dev = _abi.require_device(get_device(params.device))
    topo = mesh.edge_topology()
    if topo["boundary_edges"] or topo["nonmanifold_edges"]:
        warnings.warn(f"the mesh has {topo['boundary_edges']} boundary and {topo['nonmanifold_edges']} "
                      "non-manifold edges, so it does not bound a solid; each grid node is classified by "
                      "the majority of three axis-parallel ray casts", RuntimeWarning)
    lo, hi = mesh.bounds
    law = params.make_law(length_scale=float((hi - lo).max()))
    timings = {}
    t0 = time.time()
    grid = make_grid(mesh, params)
    tris = {}
    phi = mesh_to_sdf(mesh, grid, band=params.band, device=dev, verbose=verbose, keep_triangles=tris)
    ls = LevelSet(phi, grid, width=params.band)
    torch.cuda.current_stream(dev).synchronize()
    timings["sdf"] = time.time() - t0
    if verbose:
        print(f"[erode] {law.describe()} on a {grid.describe()}, {dev}")
    sim = ErosionSimulation(ls, law, cfl=params.cfl, feedback_cfl=params.feedback_cfl,
                            parabolic_cfl=params.parabolic_cfl, reinit_distance=params.reinit_distance,
                            reinit_iterations=params.reinit_iterations)
    if law.requires_svi and law.s == 1:
        nmax = initial_max_nr(sim)
        if nmax >= law.t:
            raise ValueError(f"t = {law.t} does not exceed N_r on the input surface (largest value "
                             f"{nmax:.4f}); the speed N^k/(t - N) needs t > N everywhere, so raise t")
        if verbose:
            print(f"[erode] initial surface: N_r <= {nmax:.4f} below t = {law.t}")
    targets = sorted({float(f) for f in params.snapshots if f > params.target_fraction},
                     reverse=True) + [params.target_fraction]
    return ErosionJob(mesh, params, law, grid, ls, sim, targets, timings, triangles=tris)


def finish_target(job: ErosionJob, stop_reason: str, fraction: float) -> bool:
    """Mesh and annotate the state reached for the current target; True if the
    job goes on to the next target (svierosion/api.py:230-245)."""
    tm = time.time()
    m = extract_surface(job.ls, project=job.params.project_vertices)
    _annotate(job.sim, m, job.mesh, job.law, job.triangles)
    job.tmesh += time.time() - tm
    job.snaps.append((fraction, job.sim.time, m))
    job.stop_reason = stop_reason
    job.fraction = fraction
    job.next_target += 1
    return stop_reason == "target_volume" and job.next_target < len(job.targets)


def job_result(job: ErosionJob, evolve_seconds: float, keep_state: bool = False) -> ErosionResult:
    job.timings["evolve"] = evolve_seconds
    job.timings["mesh"] = job.tmesh
    final = job.snaps[-1][2]
    p, law = job.params, job.law
    out = ErosionResult(final, job.snaps[:-1], job.sim.history, job.grid, job.stop_reason, job.sim.time,
                        job.fraction, job.sim.initial_volume, job.mesh.volume(), job.timings,
                        {**asdict(p), "law_description": law.describe(), **law.params()}, job.mesh)
    if keep_state:
        out.level_set, out.simulation = job.ls, job.sim
    return out


def erode(mesh: TriMesh, params: ErosionParams, callback=None, verbose: bool = False,
          keep_state: bool = False) -> ErosionResult:
    """Erode ``mesh`` under ``params.law`` until ``params.target_fraction`` of its volume remains."""
    job = prepare_job(mesh, params, verbose)
    t1 = time.time()
    while True:
        res = job.sim.run(target_fraction=job.targets[job.next_target], max_steps=params.max_steps,
                          callback=callback)
        if not finish_target(job, res.stop_reason, res.fraction):
            break
    return job_result(job, time.time() - t1 - job.tmesh, keep_state)


def volume_fraction_report(res: ErosionResult) -> str:
    """The remaining volume fraction as measured on the output mesh and on the level set."""
    before = res.initial_volume_mesh
    after = res.mesh.volume() if len(res.mesh.faces) else 0.0
    return (f"volume left: {after / before:.5f} by the output mesh ({after:.6g} of {before:.6g}), "
            f"{res.fraction:.5f} by the level set")


__all__ = ["ErosionParams", "ErosionResult", "ErosionJob", "erode", "prepare_job", "finish_target", "job_result", "make_grid", "point_mesh_distance", "volume_fraction_report",
           "StepRecord"]
