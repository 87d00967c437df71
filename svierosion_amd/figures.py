"""Mesh-to-mesh distance statistics and a cross-section figure.

``mesh_distance_stats`` is the symmetric surface distance of
svierosion/figures.py:107-114 (mean of the two directed means, RMS over both
directions, Hausdorff = the larger directed maximum), on the device distance
search of :func:`svierosion_amd.api.point_mesh_distance`.

``render_sections`` replaces the reference's off-screen renders
(svierosion/viz.py, svierosion/figures.py) with a figure that needs only
matplotlib: the three mid-plane cross sections of every mesh, overlaid, so the
eroded shapes read against the initial one.
"""

from __future__ import annotations

import numpy as np

from .meshio import TriMesh


def mesh_distance_stats(a: TriMesh, b: TriMesh, device=None) -> dict:
    """{"mean", "rms", "hausdorff"} of the vertex-to-surface distances both ways."""
    from .api import point_mesh_distance
    dab = point_mesh_distance(a.vertices, b, device=device)
    dba = point_mesh_distance(b.vertices, a, device=device)
    return {"mean": float(0.5 * (dab.mean() + dba.mean())),
            "rms": float(np.sqrt(0.5 * ((dab ** 2).mean() + (dba ** 2).mean()))),
            "hausdorff": float(max(dab.max(), dba.max()))}


def plane_section(mesh: TriMesh, axis: int, value: float) -> np.ndarray:
    """Segments (s, 2, 2) where the plane x[axis] = value cuts the triangles, in the two
    remaining coordinates (in increasing axis order)."""
    tri = mesh.vertices[mesh.faces]
    d = tri[:, :, axis] - value
    above = d > 0
    cut = above.any(1) & ~above.all(1)
    tri, d, above = tri[cut], d[cut], above[cut]
    keep = [k for k in range(3) if k != axis]
    segs = []
    for i, j in ((0, 1), (1, 2), (2, 0)):
        cross = above[:, i] != above[:, j]
        w = d[:, i] / np.where(cross, d[:, i] - d[:, j], 1.0)
        p = tri[:, i] + w[:, None] * (tri[:, j] - tri[:, i])
        segs.append((cross, p[:, keep]))
    # a cut triangle crosses exactly two of its edges: pair them in edge order
    first = np.where(segs[0][0][:, None], segs[0][1], np.where(segs[1][0][:, None], segs[1][1], segs[2][1]))
    second = np.where(segs[2][0][:, None], segs[2][1], segs[1][1])
    return np.stack([first, second], axis=1)


def render_sections(meshes: list, labels: list, path, title: str | None = None) -> None:
    """Write a PNG of the x, y and z mid-plane sections (through the first mesh's
    bounding-box centre) of every mesh, one colour per mesh."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    from matplotlib.collections import LineCollection
    lo, hi = meshes[0].bounds
    centre = 0.5 * (lo + hi)
    fig, axes = plt.subplots(1, 3, figsize=(15, 5))
    colours = plt.rcParams["axes.prop_cycle"].by_key()["color"]
    names = "xyz"
    for axis, ax in enumerate(axes):
        keep = [k for k in range(3) if k != axis]
        for n, (m, lab) in enumerate(zip(meshes, labels)):
            if len(m.faces) == 0:
                continue
            segs = plane_section(m, axis, float(centre[axis]))
            ax.add_collection(LineCollection(segs, colors=colours[n % len(colours)], linewidths=1.0,
                                             label=lab if axis == 0 else None))
        pad = 0.05 * float((hi - lo).max())
        ax.set_xlim(lo[keep[0]] - pad, hi[keep[0]] + pad)
        ax.set_ylim(lo[keep[1]] - pad, hi[keep[1]] + pad)
        ax.set_aspect("equal")
        ax.set_xlabel(names[keep[0]])
        ax.set_ylabel(names[keep[1]])
        ax.set_title(f"{names[axis]} = {centre[axis]:.4g}")
    fig.legend(loc="lower center", ncol=min(len(meshes), 7), frameon=False)
    if title:
        fig.suptitle(title)
    fig.tight_layout(rect=(0, 0.08, 1, 1))
    fig.savefig(path, dpi=120)
    plt.close(fig)
