"""Command line front end: ``python -m svierosion_amd <command> ...``.

The commands, flags, defaults, output files and exit codes are those of the
reference's ``python -m svierosion`` (svierosion/cli.py:215-259), so scripts
written against it run unchanged:

  erode     erode a PLY mesh under an SVI or curvature law to a volume fraction;
            exit status 0 when the target was reached, 2 for any other stop
  compare   several laws on one mesh to the same volume fraction (one GPU batch)
  distance  vertex-wise distances between two meshes, both directions
  shapes    write the toy solids as PLY files
  info      mesh statistics

Differences: everything computes on the ROCm GPU (``--device cpu`` is refused
with a message, there is no CPU path); ``--no-compile`` is accepted and has no
effect (the kernels are compiled ahead of time); ``--render`` and ``compare``
draw mid-plane cross sections with matplotlib instead of VTK renders.
"""

from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

PREFIX = "[svierosion]"


def _law_group(p: argparse.ArgumentParser) -> None:
    g = p.add_argument_group("flow")
    g.add_argument("--law", choices=["svi", "mean", "gauss"], default="svi",
                   help="svi: V = N^k/(t-N)^s + eps; mean: V = (H_+)^alpha; gauss: V = (K_+)^alpha")
    g.add_argument("--radius", "-r", type=float, help="ball radius r of the SVI law, in mesh units")
    g.add_argument("--radius-rel", type=float, help="ball radius as a fraction of the longest bounding-box side")
    g.add_argument("-k", type=int, default=1, help="SVI exponent k (default 1)")
    g.add_argument("-s", type=int, default=0, choices=[0, 1], help="SVI singular term on (1) or off (0)")
    g.add_argument("-t", type=float, help="SVI singularity t, required with -s 1")
    g.add_argument("--eps", type=float, default=0.0, help="constant speed added to the law")
    g.add_argument("--alpha", type=float, default=1.0, help="exponent of the curvature laws")
    g.add_argument("--delta", type=float, default=None,
                   help="regularisation of x^alpha at 0 when alpha < 1 (default from the grid)")
    g.add_argument("--max-speed-ratio", type=float, default=100.0,
                   help="speed cap of the s = 1 SVI law, relative to the flat-surface speed")


def _numerics_group(p: argparse.ArgumentParser) -> None:
    g = p.add_argument_group("numerics")
    g.add_argument("--target", type=float, default=0.9, help="volume fraction to stop at (default 0.9)")
    g.add_argument("--snapshots", type=float, nargs="*", default=[],
                   help="further volume fractions at which a mesh is written")
    g.add_argument("--spacing", type=float, help="grid spacing h in mesh units")
    g.add_argument("--resolution", type=int, default=200,
                   help="cells along the longest bounding-box side when --spacing is not given")
    g.add_argument("--band", type=float, default=6.0, help="narrow-band half width in cells")
    g.add_argument("--cfl", type=float, default=0.5)
    g.add_argument("--device", default=None, help="cuda or cuda:N (default: the first GPU)")
    g.add_argument("--no-compile", action="store_true", help="accepted; the kernels are prebuilt")
    g.add_argument("--no-project", action="store_true",
                   help="leave mesh vertices on the piecewise-linear zero set (no tricubic projection)")


def params_from_args(args, mesh):
    """ErosionParams of an ``erode`` command line (svierosion/cli.py:59-71): an SVI run
    without --radius takes --radius-rel times the longest bounding-box side."""
    from .api import ErosionParams
    radius = args.radius
    if args.law == "svi" and radius is None:
        if args.radius_rel is None:
            raise SystemExit("error: the SVI law needs --radius or --radius-rel")
        lo, hi = mesh.bounds
        radius = args.radius_rel * float((hi - lo).max())
    return ErosionParams(law=args.law, radius=radius, k=args.k, s=args.s, t=args.t, eps=args.eps,
                         max_speed_ratio=args.max_speed_ratio, alpha=args.alpha, delta=args.delta,
                         target_fraction=args.target, snapshots=tuple(args.snapshots), spacing=args.spacing,
                         resolution=args.resolution, band=args.band, cfl=args.cfl, device=args.device,
                         project_vertices=not args.no_project)


def progress_printer(enabled: bool):
    """Step callback printing the first step and then at most one line every 2 s."""
    last = [time.perf_counter()]

    def show(sim, rec):
        now = time.perf_counter()
        if not enabled or (rec.step != 1 and now - last[0] <= 2.0):
            return
        last[0] = now
        q = "N" if sim.law.requires_svi else "H"
        fmt = ".3f" if sim.law.requires_svi else ".3g"
        line = (f"  step {rec.step:6d}  t={rec.time:.5g}  dt={rec.dt:.2e}  V/V0={rec.fraction:.5f}  "
                f"vmax={rec.vmax:.3g} {q}=[{rec.qmin:{fmt}},{rec.qmax:{fmt}}]")
        if rec.capped:
            line += f" capped={rec.capped}"
        print(line, flush=True)
    return show


def _check_device(device) -> None:
    from . import _abi
    from .grid import get_device
    try:
        _abi.require_device(get_device(device))
    except RuntimeError as e:
        raise SystemExit(f"error: {e}")


def cmd_erode(args) -> int:
    from .api import erode, volume_fraction_report
    from .plyio import read_ply
    from .results import write_result
    _check_device(args.device)
    mesh = read_ply(args.input)
    params = params_from_args(args, mesh)
    src = Path(args.input)
    out = Path(args.output) if args.output else src.with_name(f"{src.stem}_eroded.ply")
    say = (lambda s: None) if args.quiet else (lambda s: print(f"{PREFIX} {s}", flush=True))
    say(f"{args.input}: {len(mesh.vertices)} vertices, {len(mesh.faces)} triangles, volume {mesh.volume():.6g}, "
        f"watertight {mesh.is_watertight()}")
    res = erode(mesh, params, callback=progress_printer(not args.quiet), verbose=not args.quiet)
    written = write_result(res, out, args.quality)
    say(res.summary())
    say(volume_fraction_report(res))
    for p in written:
        say(f"wrote {p}")
    if args.render:
        from .figures import render_sections
        render_sections([mesh, res.mesh], ["initial", f"eroded ({res.fraction:.3f})"], args.render)
        say(f"wrote {args.render}")
    return 0 if res.stop_reason == "target_volume" else 2


def cmd_compare(args) -> int:
    from .compare import default_comparison_laws, run_comparison
    from .plyio import read_ply
    _check_device(args.device)
    mesh = read_ply(args.input)
    if args.radius is not None:
        radius = args.radius
    else:
        lo, hi = mesh.bounds
        radius = (args.radius_rel or 0.15) * float((hi - lo).max())
    run_comparison(mesh, default_comparison_laws(radius, t=args.t), Path(args.output), target=args.target,
                   resolution=args.resolution, spacing=args.spacing, device=args.device, verbose=not args.quiet)
    return 0


def _stats(d: np.ndarray) -> str:
    return f"mean {d.mean():.6g}  rms {np.sqrt((d * d).mean()):.6g}  max {d.max():.6g}"


def cmd_distance(args) -> int:
    from .api import point_mesh_distance
    from .plyio import read_ply, write_ply
    _check_device(args.device)
    a, b = read_ply(args.mesh_a), read_ply(args.mesh_b)
    ab = point_mesh_distance(a.vertices, b, device=args.device)
    ba = point_mesh_distance(b.vertices, a, device=args.device)
    print(f"A->B: {_stats(ab)}")
    print(f"B->A: {_stats(ba)}")
    print(f"Hausdorff {max(ab.max(), ba.max()):.6g}")
    if args.output:
        write_ply(args.output, a, vertex_data={"quality": ab.astype(np.float32)})
        print(f"wrote {args.output} (vertex quality = distance to {args.mesh_b})")
    return 0


def cmd_shapes(args) -> int:
    from .plyio import write_ply
    from .shapes import SHAPES
    out = Path(args.output)
    out.mkdir(parents=True, exist_ok=True)
    for name, make in SHAPES.items():
        m = make()
        path = out / f"{name}.ply"
        write_ply(path, m)
        print(f"{name:14s} {len(m.faces):7d} triangles  volume {m.volume():.5f}  -> {path}")
    return 0


def cmd_info(args) -> int:
    from .plyio import read_ply
    m = read_ply(args.input)
    lo, hi = m.bounds
    topo = m.edge_topology()
    print(f"vertices {len(m.vertices)}, triangles {len(m.faces)}")
    print(f"bounding box {lo} .. {hi} (extent {hi - lo})")
    print(f"area {m.area():.6g}, signed volume {m.signed_volume():.6g}")
    print(f"edges {topo['edges']}, boundary {topo['boundary_edges']}, non-manifold {topo['nonmanifold_edges']}, "
          f"misoriented {topo['misoriented_edges']}")
    print(f"vertex properties: {sorted(m.vertex_data)}")
    return 0


def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="svierosion_amd", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="command", required=True)

    p = sub.add_parser("erode", help="erode a mesh")
    p.add_argument("input")
    p.add_argument("-o", "--output", help="output PLY (default <input>_eroded.ply)")
    _law_group(p)
    _numerics_group(p)
    p.add_argument("--quality", default="depth",
                   help="vertex field copied to 'quality' for MeshLab colouring: depth, N, speed, H, K")
    p.add_argument("--render", help="also write a PNG of initial and eroded cross sections")
    p.add_argument("--quiet", action="store_true")
    p.set_defaults(func=cmd_erode)

    p = sub.add_parser("compare", help="SVI and curvature flows on one mesh")
    p.add_argument("input")
    p.add_argument("-o", "--output", required=True, help="output directory")
    p.add_argument("--radius", "-r", type=float)
    p.add_argument("--radius-rel", type=float, help="radius over the longest bounding-box side (default 0.15)")
    p.add_argument("-t", type=float, default=0.93, help="t of the s = 1 SVI law")
    p.add_argument("--target", type=float, default=0.9)
    p.add_argument("--spacing", type=float)
    p.add_argument("--resolution", type=int, default=160)
    p.add_argument("--device", default=None)
    p.add_argument("--no-compile", action="store_true")
    p.add_argument("--quiet", action="store_true")
    p.set_defaults(func=cmd_compare)

    p = sub.add_parser("distance", help="distances between two meshes")
    p.add_argument("mesh_a")
    p.add_argument("mesh_b")
    p.add_argument("-o", "--output", help="write mesh A with its distance to B as vertex quality")
    p.add_argument("--device", default=None)
    p.set_defaults(func=cmd_distance)

    p = sub.add_parser("shapes", help="write the toy solids")
    p.add_argument("output")
    p.set_defaults(func=cmd_shapes)

    p = sub.add_parser("info", help="mesh statistics")
    p.add_argument("input")
    p.set_defaults(func=cmd_info)
    return ap


def main(argv=None) -> int:
    args = build_parser().parse_args(argv)
    return args.func(args)


if __name__ == "__main__":  # pragma: no cover
    sys.exit(main())
