"""Several erosion laws on one solid, all stopped at the same remaining volume.

Same outputs as svierosion/compare.py:36-85 -- ``initial.ply``, one PLY per law
(vertex ``quality`` = erosion depth), ``summary.csv`` (one row per law),
``pairwise_mean_distance.json`` (symmetric mean surface distance between the
final shapes) and a figure -- but the laws run as one batch on the GPU
(:func:`svierosion_amd.batch.serve`): every law is a job of the batched
engine, so the comparison takes about as long as its slowest law instead of
the sum.  Each job's results are those of ``erode`` on its own (the batched
and single-job paths share the kernels and the step order).
"""

from __future__ import annotations

import csv
import json
import re
from pathlib import Path

import numpy as np

from .api import ErosionParams
from .meshio import TriMesh
from .plyio import write_ply


def default_comparison_laws(radius: float, t: float = 0.93) -> list:
    """(label, ErosionParams keywords): grassfire, SVI N, N^4 and N/(t-N), mean and
    Gauss curvature (the set of svierosion/compare.py:18-29)."""
    svi = dict(law="svi", radius=radius)
    return [("grassfire V=1", dict(svi, k=0, s=0)),
            ("SVI V=N", dict(svi, k=1, s=0)),
            ("SVI V=N^4", dict(svi, k=4, s=0)),
            (f"SVI V=N/({t:g}-N)", dict(svi, k=1, s=1, t=t)),
            ("mean curv. V=H+", dict(law="mean", alpha=1.0)),
            ("Gauss curv. V=K+", dict(law="gauss", alpha=1.0))]


def slug(label: str) -> str:
    """File stem of a law label: runs of non-alphanumerics become one underscore, lower case."""
    return "_".join(w for w in re.split(r"[^A-Za-z0-9]+", label) if w).lower()


def _row(label: str, res, initial_volume: float, wall: float) -> dict:
    depth = res.mesh.vertex_data.get("depth")
    if depth is None or len(depth) == 0:
        depth = np.zeros(1)
    mesh_fraction = res.mesh.volume() / initial_volume if len(res.mesh.faces) else 0.0
    return {"label": label, "stop": res.stop_reason, "steps": len(res.history), "model_time": res.time,
            "volume_fraction": res.fraction, "mesh_volume_fraction": mesh_fraction,
            "depth_mean": float(depth.mean()), "depth_max": float(depth.max()),
            "wall_seconds": wall, "grid": "x".join(str(n) for n in res.grid.shape)}


def run_comparison(mesh: TriMesh, laws: list, outdir, target: float = 0.9, resolution: int = 160,
                   spacing: float | None = None, device=None, verbose: bool = True, max_steps: int = 200_000,
                   title: str | None = None, figure: bool = True) -> list:
    """Erode ``mesh`` under every (label, keywords) of ``laws`` and write the outputs to
    ``outdir``; returns [(label, ErosionResult)].  A law that raises stops the comparison
    with that error.  ``wall_seconds`` is a law's time from admission to completion in
    the batch."""
    from .batch import Request, serve
    from .figures import mesh_distance_stats, render_sections
    from .grid import get_device
    from . import _abi
    _abi.require_device(get_device(device))
    outdir = Path(outdir)
    outdir.mkdir(parents=True, exist_ok=True)
    write_ply(outdir / "initial.ply", mesh)
    requests = [Request(mesh, ErosionParams(**kw, target_fraction=target, resolution=resolution, spacing=spacing,
                                            device=device, max_steps=max_steps)) for _, kw in laws]
    report = serve(requests, max_jobs=len(requests))
    v0 = mesh.volume()
    results, rows = [], []
    for (label, _), c in zip(laws, report.completions):
        if c.error is not None:
            raise RuntimeError(f"{label}: {c.error}")
        res = c.result
        depth = res.mesh.vertex_data.get("depth", np.zeros(len(res.mesh.vertices)))
        write_ply(outdir / f"{slug(label)}.ply", res.mesh, vertex_data={"quality": depth})
        row = _row(label, res, v0, c.finished - c.admitted)
        rows.append(row)
        results.append((label, res))
        if verbose:
            print(f"[compare] {label:24s} {res.stop_reason:13s} steps {row['steps']:6d}  "
                  f"V/V0 {row['volume_fraction']:.4f}  depth mean {row['depth_mean']:.4g} "
                  f"max {row['depth_max']:.4g}  ({row['wall_seconds']:.1f}s)", flush=True)
    with open(outdir / "summary.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)
    finals = [r.mesh for _, r in results]
    n = len(finals)
    dist = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            if len(finals[i].faces) and len(finals[j].faces):
                dist[i, j] = dist[j, i] = mesh_distance_stats(finals[i], finals[j], device=device)["mean"]
    labels = [lab for lab, _ in results]
    (outdir / "pairwise_mean_distance.json").write_text(
        json.dumps({"labels": labels, "mean_distance": dist.tolist()}, indent=2) + "\n")
    if figure:
        render_sections([mesh] + finals, ["initial"] + labels, outdir / "comparison.png",
                        title=title or f"all runs stopped at V/V0 = {target:g}")
    return results
