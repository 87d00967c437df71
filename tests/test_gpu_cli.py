"""The command line front end on the GPU: ``erode`` writes what the reference's
writes (svierosion/cli.py:91-128) with the results of ``erode()``, exits 0 at
the target and 2 otherwise (svierosion/cli.py:150), ``distance`` and
``compare`` produce their reports (svierosion/cli.py:169-182,
svierosion/compare.py:36-85)."""

import csv
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _write(path, mesh):
    from svierosion_amd.plyio import write_ply
    write_ply(path, mesh)
    return str(path)


def test_erode_command_writes_the_result_of_erode(device, tmp_path, capsys):
    from svierosion_amd import ErosionParams, cli, erode
    from svierosion_amd.plyio import read_ply
    from svierosion_amd.results import read_history
    from svierosion_amd.shapes import cube
    src = _write(tmp_path / "cube.ply", cube(1.0))
    rc = cli.main(["erode", src, "-r", "0.2", "-k", "1", "--resolution", "40", "--target", "0.94",
                   "--snapshots", "0.97"])
    out = capsys.readouterr().out
    print(out[-600:])
    assert rc == 0
    assert "target_volume" in out and "volume left" in out
    ref = erode(read_ply(src), ErosionParams(law="svi", radius=0.2, k=1, resolution=40, target_fraction=0.94,
                                             snapshots=(0.97,)))
    final = read_ply(tmp_path / "cube_eroded.ply")
    assert final.faces.shape == ref.mesh.faces.shape
    assert np.array_equal(final.vertices, ref.mesh.vertices.astype(np.float32).astype(np.float64))
    assert np.array_equal(final.vertex_data["quality"], ref.mesh.vertex_data["depth"].astype(np.float32))
    assert (tmp_path / "cube_eroded_v0.970.ply").exists()
    hist = read_history(tmp_path / "cube_eroded_history.csv")
    assert [r.dt for r in hist] == [r.dt for r in ref.history]
    assert [r.fraction for r in hist] == [r.fraction for r in ref.history]
    meta = json.loads((tmp_path / "cube_eroded.json").read_text())
    assert meta["stop_reason"] == "target_volume" and meta["time"] == ref.time
    assert meta["grid"]["shape"] == list(ref.grid.shape)
    assert [s["fraction"] for s in meta["snapshots"]] == [f for f, _, _ in ref.snapshots]


def test_erode_command_exits_2_when_the_target_is_not_reached(device, tmp_path, monkeypatch):
    from svierosion_amd import cli
    from svierosion_amd.shapes import cube
    src = _write(tmp_path / "cube.ply", cube(1.0))
    base = cli.params_from_args

    def three_steps(args, mesh):
        p = base(args, mesh)
        p.max_steps = 3
        return p
    monkeypatch.setattr(cli, "params_from_args", three_steps)
    rc = cli.main(["erode", src, "-r", "0.2", "--resolution", "32", "--target", "0.5", "--quiet",
                   "-o", str(tmp_path / "o.ply")])
    assert rc == 2
    meta = json.loads((tmp_path / "o.json").read_text())
    assert meta["stop_reason"] == "max_steps"
    assert len(list(csv.DictReader(open(tmp_path / "o_history.csv")))) == 3


def test_distance_command_between_concentric_spheres(device, tmp_path, capsys):
    from svierosion_amd import cli
    from svierosion_amd.plyio import read_ply
    from svierosion_amd.shapes import icosphere
    a = _write(tmp_path / "a.ply", icosphere(4, 1.0))
    b = _write(tmp_path / "b.ply", icosphere(4, 0.9))
    assert cli.main(["distance", a, b, "-o", str(tmp_path / "d.ply")]) == 0
    out = capsys.readouterr().out
    print(out)
    # every vertex of A lies 0.1 outside B's vertices; B's flat facets are at most the
    # sagitta of a level-4 icosphere (< 1.5e-3 of the radius) inside the sphere
    d = read_ply(tmp_path / "d.ply").vertex_data["quality"]
    assert np.all(d >= 0.1 - 1e-5) and np.all(d <= 0.1 + 1.5e-3)
    haus = float(out.split("Hausdorff")[1].split()[0])
    assert 0.1 - 1e-5 <= haus <= 0.1 + 1.5e-3


def test_compare_command_batches_the_default_laws(device, tmp_path, capsys):
    from svierosion_amd import ErosionParams, cli, erode
    from svierosion_amd.compare import default_comparison_laws
    from svierosion_amd.shapes import icosphere
    src = _write(tmp_path / "s.ply", icosphere(3, 1.0))
    outdir = tmp_path / "cmp"
    assert cli.main(["compare", src, "-o", str(outdir), "--resolution", "32", "--target", "0.96", "-t", "0.9",
                     "--quiet"]) == 0
    rows = list(csv.DictReader(open(outdir / "summary.csv")))
    laws = default_comparison_laws(0.15 * 2.0, t=0.9)
    assert [r["label"] for r in rows] == [lab for lab, _ in laws]
    assert all(r["stop"] == "target_volume" for r in rows)
    for r in rows:
        assert abs(float(r["volume_fraction"]) - 0.96) < 1e-5
    pair = json.loads((outdir / "pairwise_mean_distance.json").read_text())
    D = np.array(pair["mean_distance"])
    assert D.shape == (6, 6) and np.allclose(D, D.T) and np.all(np.diag(D) == 0)
    assert (outdir / "comparison.png").stat().st_size > 0
    for name in ("initial", "grassfire_v_1", "svi_v_n", "gauss_curv_v_k"):
        assert (outdir / f"{name}.ply").exists()
    # the batched run of a law equals erode() of that law alone
    from svierosion_amd.plyio import read_ply
    label, kw = laws[1]
    alone = erode(read_ply(src), ErosionParams(**kw, target_fraction=0.96, resolution=32))
    assert float(rows[1]["model_time"]) == alone.time
    assert int(rows[1]["steps"]) == len(alone.history)
