"""On-disk results of an erosion: the final and snapshot meshes (PLY), the
per-step history (CSV) and a JSON record of parameters, grid, timings and
snapshots.  File names and contents follow the reference's writers
(svierosion/cli.py:91-128):

* ``<out>``                   final mesh, with the ``quality`` vertex scalar;
* ``<stem>_v<fraction>.ply``  one mesh per snapshot (fraction to 3 decimals);
* ``<stem>_history.csv``      one row per accepted step (StepRecord fields);
* ``<stem>.json``             summary, parameters, stop reason, grid, timings,
                              mesh volumes and snapshot times.
"""

from __future__ import annotations

import csv
import json
from dataclasses import asdict, fields
from pathlib import Path

from .evolve import StepRecord
from .plyio import write_ply


def write_history(path, history) -> None:
    """One CSV row per StepRecord, header = the record's field names
    (svierosion/cli.py:91-98); nothing is written for an empty history."""
    if not history:
        return
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=list(asdict(history[0])))
        w.writeheader()
        for rec in history:
            w.writerow(asdict(rec))


def read_history(path) -> list:
    """StepRecords back from a history CSV written by :func:`write_history`."""
    types = {f.name: f.type for f in fields(StepRecord)}
    out = []
    with open(path, newline="") as fh:
        for row in csv.DictReader(fh):
            rec = {}
            for k, v in row.items():
                t = types.get(k, "float")
                if t in ("int", int):
                    rec[k] = int(v)
                elif t in ("bool", bool):
                    rec[k] = v == "True"
                else:
                    rec[k] = float(v)
            out.append(StepRecord(**rec))
    return out


def _write_mesh(path, mesh, quality: str | None) -> None:
    extra = {}
    if quality and quality in mesh.vertex_data:
        extra["quality"] = mesh.vertex_data[quality]
    write_ply(path, mesh, vertex_data=extra)


def result_meta(res) -> dict:
    """The JSON record of an ErosionResult (svierosion/cli.py:111-119)."""
    return {"summary": res.summary(), "params": res.params, "stop_reason": res.stop_reason,
            "time": res.time, "volume_fraction": res.fraction,
            "grid": {"origin": list(res.grid.origin), "spacing": res.grid.spacing, "shape": list(res.grid.shape)},
            "timings": res.timings,
            "initial_mesh_volume": res.initial_volume_mesh,
            "final_mesh_volume": res.mesh.volume() if len(res.mesh.faces) else 0.0,
            "snapshots": [{"fraction": f, "time": t} for f, t, _ in res.snapshots]}


def write_result(res, out, quality: str | None = "depth") -> list:
    """Write every file of ``res`` next to ``out`` (the final mesh's path);
    returns the mesh paths written, snapshots first."""
    out = Path(out)
    stem = out.with_suffix("")
    written = []
    for frac, _, m in res.snapshots:
        p = Path(f"{stem}_v{frac:.3f}.ply")
        _write_mesh(p, m, quality)
        written.append(p)
    _write_mesh(out, res.mesh, quality)
    written.append(out)
    write_history(Path(f"{stem}_history.csv"), res.history)
    Path(f"{stem}.json").write_text(json.dumps(result_meta(res), indent=2, default=float) + "\n")
    return written
