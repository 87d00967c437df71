"""Command line front end and the toy solids, on the CPU.

Mirrors the reference's CLI surface (svierosion/cli.py:215-259): flags and
defaults, the flag -> ErosionParams mapping, the ``shapes`` and ``info``
commands, and the refusal of a CPU device (the compute commands run on the GPU
only; their GPU test is tests/test_gpu_cli.py).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from svierosion_amd import cli, shapes
from svierosion_amd.compare import default_comparison_laws, slug
from svierosion_amd.figures import plane_section
from svierosion_amd.plyio import read_ply


def _erode_args(*extra):
    return cli.build_parser().parse_args(["erode", "in.ply", *extra])


def test_erode_defaults_match_the_reference_flags():
    a = _erode_args()
    assert (a.law, a.k, a.s, a.t, a.eps, a.alpha, a.delta, a.max_speed_ratio) == ("svi", 1, 0, None, 0.0, 1.0, None, 100.0)
    assert (a.target, a.snapshots, a.spacing, a.resolution, a.band, a.cfl) == (0.9, [], None, 200, 6.0, 0.5)
    assert (a.quality, a.render, a.quiet, a.no_compile, a.no_project) == ("depth", None, False, False, False)
    c = cli.build_parser().parse_args(["compare", "in.ply", "-o", "out"])
    assert (c.t, c.target, c.resolution, c.radius, c.radius_rel) == (0.93, 0.9, 160, None, None)
    with pytest.raises(SystemExit):
        cli.build_parser().parse_args(["erode", "in.ply", "-s", "2"])
    with pytest.raises(SystemExit):
        cli.build_parser().parse_args(["compare", "in.ply"])       # -o is required


def test_params_from_args_maps_every_flag():
    mesh = shapes.box((2.0, 1.0, 0.5))
    a = _erode_args("--radius-rel", "0.25", "-k", "2", "-s", "1", "-t", "0.8", "--eps", "0.1",
                    "--max-speed-ratio", "50", "--target", "0.7", "--snapshots", "0.95", "0.8",
                    "--resolution", "96", "--band", "5", "--cfl", "0.4", "--no-project")
    p = cli.params_from_args(a, mesh)
    assert p.radius == pytest.approx(0.5)           # 0.25 x the longest side (2.0)
    assert (p.k, p.s, p.t, p.eps, p.max_speed_ratio) == (2, 1, 0.8, 0.1, 50.0)
    assert (p.target_fraction, p.snapshots, p.resolution, p.band, p.cfl) == (0.7, (0.95, 0.8), 96, 5.0, 0.4)
    assert p.project_vertices is False and p.spacing is None
    p = cli.params_from_args(_erode_args("-r", "0.3", "--radius-rel", "0.9", "--spacing", "0.01"), mesh)
    assert p.radius == 0.3 and p.spacing == 0.01    # an absolute radius wins
    p = cli.params_from_args(_erode_args("--law", "gauss", "--alpha", "0.25", "--delta", "2"), mesh)
    assert (p.law, p.alpha, p.delta, p.radius) == ("gauss", 0.25, 2.0, None)
    with pytest.raises(SystemExit, match="needs --radius"):
        cli.params_from_args(_erode_args(), mesh)


def test_compute_commands_refuse_the_cpu(tmp_path, capsys):
    path = tmp_path / "cube.ply"
    from svierosion_amd.plyio import write_ply
    write_ply(path, shapes.cube())
    for argv in (["erode", str(path), "-r", "0.1", "--device", "cpu"],
                 ["distance", str(path), str(path), "--device", "cpu"],
                 ["compare", str(path), "-o", str(tmp_path / "cmp"), "--device", "cpu"]):
        with pytest.raises(SystemExit, match="no CPU fallback"):
            cli.main(argv)


def _polygon_area(n, r):
    return 0.5 * n * r * r * math.sin(2 * math.pi / n)


EXPECTED_VOLUMES = {
    "cube": 1.0,
    "box": 1.6 * 1.0 * 0.6,
    "prism": math.sqrt(3) / 4 * 1.2 ** 2 * 1.2,
    "lblock": (1.2 ** 2 - 0.6 ** 2) * 0.6,
    "cylinder": _polygon_area(256, 0.45) * 1.2,
    # annular sector as two polygonal arcs of 160 chords over 2.2 rad
    "tube_fragment": 0.5 * 160 * math.sin(2.2 / 160) * (1.0 - 0.36) * 1.6,
}


def test_shapes_command_writes_closed_outward_solids(tmp_path, capsys):
    assert cli.main(["shapes", str(tmp_path)]) == 0
    lines = capsys.readouterr().out.strip().splitlines()
    assert [ln.split()[0] for ln in lines] == list(shapes.SHAPES)
    for name in shapes.SHAPES:
        m = read_ply(tmp_path / f"{name}.ply")
        topo = m.edge_topology()
        assert topo["boundary_edges"] == 0 and topo["nonmanifold_edges"] == 0 and topo["misoriented_edges"] == 0, name
        assert m.signed_volume() > 0, name
        if name in EXPECTED_VOLUMES:
            # exact in memory; the PLY file holds float32 vertices
            assert shapes.SHAPES[name]().volume() == pytest.approx(EXPECTED_VOLUMES[name], rel=1e-12), name
            assert m.volume() == pytest.approx(EXPECTED_VOLUMES[name], rel=1e-6), name
    # smooth solids converge to their analytic volumes
    sphere = read_ply(tmp_path / "sphere.ply")
    assert sphere.volume() == pytest.approx(4 / 3 * math.pi * 0.6 ** 3, rel=2e-3)
    ell = read_ply(tmp_path / "ellipsoid.ply")
    assert ell.volume() == pytest.approx(4 / 3 * math.pi * 0.8 * 0.55 * 0.4, rel=2e-3)
    torus = read_ply(tmp_path / "torus.ply")
    assert torus.volume() == pytest.approx(2 * math.pi ** 2 * 0.6 * 0.25 ** 2, rel=2e-3)
    assert len(torus.faces) == 2 * 192 * 96


def test_prism_caps_have_outward_normals():
    for m in (shapes.l_block(1.2, 0.6, 0.6), shapes.tube_fragment(), shapes.triangular_prism(1.2, 1.2)):
        n = m.face_normals()
        c = m.triangles().mean(1)
        top = np.isclose(c[:, 2], m.vertices[:, 2].max())
        bottom = np.isclose(c[:, 2], m.vertices[:, 2].min())
        assert top.any() and bottom.any()
        assert np.allclose(n[top], [0, 0, 1]) and np.allclose(n[bottom], [0, 0, -1])


def test_analytic_sdfs():
    p = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.2, 0.1, 0.0]])
    assert np.allclose(shapes.sphere_sdf(p, radius=0.5), np.linalg.norm(p, axis=1) - 0.5)
    d = shapes.box_sdf(p, size=(1.0, 1.0, 1.0))
    assert np.allclose(d, [-0.5, 0.5, 0.0, -0.3])
    corner = shapes.box_sdf(np.array([[1.5, 1.5, 0.0]]))
    assert np.allclose(corner, math.sqrt(2.0))


def test_info_command(tmp_path, capsys):
    from svierosion_amd.plyio import write_ply
    path = tmp_path / "l.ply"
    write_ply(path, shapes.l_block(1.2, 0.6, 0.6), vertex_data={"quality": np.zeros(12)})
    assert cli.main(["info", str(path)]) == 0
    out = capsys.readouterr().out
    assert "vertices 12, triangles 20" in out
    assert "signed volume 0.648" in out
    assert "boundary 0, non-manifold 0, misoriented 0" in out
    assert "vertex properties: ['quality']" in out


def test_comparison_laws_and_slugs():
    laws = default_comparison_laws(0.2, t=0.93)
    assert [lab for lab, _ in laws] == ["grassfire V=1", "SVI V=N", "SVI V=N^4", "SVI V=N/(0.93-N)",
                                        "mean curv. V=H+", "Gauss curv. V=K+"]
    assert [slug(lab) for lab, _ in laws] == ["grassfire_v_1", "svi_v_n", "svi_v_n_4", "svi_v_n_0_93_n",
                                              "mean_curv_v_h", "gauss_curv_v_k"]
    assert laws[3][1] == dict(law="svi", radius=0.2, k=1, s=1, t=0.93)


def test_plane_section_of_a_cube_is_its_square():
    segs = plane_section(shapes.cube(1.0), 2, 0.1)
    length = np.linalg.norm(segs[:, 1] - segs[:, 0], axis=1).sum()
    assert length == pytest.approx(4.0)
    assert np.allclose(np.abs(segs).max(), 0.5)
