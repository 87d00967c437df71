"""Host-side pieces of the drop-in API that need no GPU: grid sizing, meshes,
PLY round trips (reference tests/test_geometry.py:20-49), generators, law parameters,
surface extraction topology."""

import os
import tempfile

import numpy as np
import pytest

from oracle import stencils as S


def test_grid_matches_oracle_sizing():
    from svierosion_amd.api import ErosionParams, make_grid
    from svierosion_amd import shapes
    m = shapes.cube(1.0)
    g = make_grid(m, ErosionParams(law="svi", radius=0.1, resolution=64))
    assert g.shape == (90, 90, 90) and g.spacing == 1 / 64
    assert np.allclose(g.origin, -0.6953125)
    o, h, shp = S.make_grid(*m.bounds, resolution=64, radius=0.1)
    assert shp == g.shape and np.allclose(o, g.origin)


def test_ply_roundtrip_binary_and_ascii():
    from svierosion_amd import shapes
    from svierosion_amd.plyio import read_ply, write_ply
    m = shapes.torus(1.0, 0.3, 32, 16)
    q = m.vertices[:, 2].copy()
    with tempfile.TemporaryDirectory() as d:
        for binary, prec in ((True, "float"), (True, "double"), (False, "double")):
            p = os.path.join(d, f"t_{binary}_{prec}.ply")
            write_ply(p, m, binary=binary, precision=prec, vertex_data={"quality": q})
            r = read_ply(p)
            tol = 1e-6 if prec == "float" else 1e-12
            assert np.allclose(r.vertices, m.vertices, atol=tol)
            assert np.array_equal(r.faces, m.faces)
            assert np.allclose(r.vertex_data["quality"], q, atol=1e-6)


def test_reference_import_paths():
    """The reference's import paths: read_ply / write_ply from the package and from
    its meshio module (svierosion/__init__.py:25, svierosion/meshio.py:264,313)."""
    import svierosion_amd
    from svierosion_amd import meshio, plyio
    from svierosion_amd.meshio import read_ply, write_ply
    assert read_ply is plyio.read_ply and write_ply is plyio.write_ply
    assert svierosion_amd.read_ply is read_ply and svierosion_amd.TriMesh is meshio.TriMesh
    with pytest.raises(AttributeError):
        meshio.not_a_function


def test_ply_quads_are_triangulated():
    from svierosion_amd.plyio import read_ply
    txt = ("ply\nformat ascii 1.0\nelement vertex 4\nproperty float x\nproperty float y\n"
           "property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n"
           "0 0 0\n1 0 0\n1 1 0\n0 1 0\n4 0 1 2 3\n")
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "q.ply")
        open(p, "w").write(txt)
        assert read_ply(p).faces.shape == (2, 3)


def test_ply_vertex_properties_flag_types_and_writer_options():
    """read_ply(vertex_properties=True/False) as in svierosion/meshio.py:264 (a list of
    names is an extension), extra properties keep their on-disk types (binary and
    ascii), write_ply creates missing directories and writes comment lines
    (svierosion/meshio.py:313-333)."""
    from svierosion_amd import shapes
    from svierosion_amd.plyio import read_ply, write_ply
    m = shapes.icosphere(1)
    q = np.linspace(0, 1, len(m.vertices))
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "new", "deeper", "s.ply")
        write_ply(p, m, vertex_data={"quality": q, "depth": 2 * q}, comments=["made by a test", "two\nlines"])
        head = open(p, "rb").read(400).split(b"end_header")[0].decode()
        assert "comment made by a test\n" in head and "comment two lines\n" in head
        a = read_ply(p, vertex_properties=True)
        assert set(a.vertex_data) == {"quality", "depth"} and a.vertex_data["quality"].dtype == np.float32
        assert read_ply(p, vertex_properties=False).vertex_data == {}
        assert set(read_ply(p, vertex_properties=["depth"]).vertex_data) == {"depth"}
        for fmt, rows in (("binary_big_endian", None), ("ascii", None)):
            txt = (f"ply\nformat {fmt} 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                   "property float z\nproperty uchar label\nproperty double w\nelement face 1\n"
                   "property list uchar int vertex_indices\nend_header\n").encode()
            if fmt == "ascii":
                body = b"0 0 0 7 0.25\n1 0 0 8 0.5\n0 1 0 9 0.75\n3 0 1 2\n"
            else:
                vdt = np.dtype([("x", ">f4"), ("y", ">f4"), ("z", ">f4"), ("label", "u1"), ("w", ">f8")])
                va = np.array([(0, 0, 0, 7, .25), (1, 0, 0, 8, .5), (0, 1, 0, 9, .75)], dtype=vdt)
                body = va.tobytes() + np.array([3], "u1").tobytes() + np.array([0, 1, 2], ">i4").tobytes()
            f = os.path.join(d, f"t_{fmt}.ply")
            open(f, "wb").write(txt + body)
            r = read_ply(f)
            assert r.vertex_data["label"].dtype == np.uint8 and r.vertex_data["label"].tolist() == [7, 8, 9]
            assert r.vertex_data["w"].dtype == np.float64 and r.vertex_data["w"].tolist() == [.25, .5, .75]
            assert r.faces.tolist() == [[0, 1, 2]]
        with pytest.raises(ValueError):
            write_ply(os.path.join(d, "bad.ply"), m, vertex_data={"x2": q[:-1]})


@pytest.mark.parametrize("name", ["cube", "icosphere", "torus", "sh_fragment", "noisy_ellipsoid"])
def test_generators_closed_and_outward(name):
    from svierosion_amd import shapes
    make = {"cube": lambda: shapes.cube(1.0), "icosphere": lambda: shapes.icosphere(3),
            "torus": lambda: shapes.torus(1.0, 0.35, 32, 16), "sh_fragment": lambda: shapes.sh_fragment(3, seed=1),
            "noisy_ellipsoid": lambda: shapes.noisy_ellipsoid(3, seed=2)}[name]
    m = make()
    assert m.is_watertight() and m.is_consistently_oriented() and m.signed_volume() > 0


def test_icosphere_counts():
    from svierosion_amd import shapes
    for s in range(4):
        m = shapes.icosphere(s)
        assert len(m.faces) == 20 * 4 ** s
        assert np.allclose(np.linalg.norm(m.vertices, axis=1), 1.0)


def test_law_parameters_and_errors():
    from svierosion_amd.api import ErosionParams
    from svierosion_amd.flows import SVIPowerLaw, flat_speed
    law = SVIPowerLaw(0.3, k=2, s=1, t=0.93)
    assert abs(law._raw(law.n_cap) / flat_speed(law) - 100) < 1e-6
    for bad in (dict(radius=-1), dict(radius=1, k=-1), dict(radius=1, s=2), dict(radius=1, s=1, t=0.4)):
        with pytest.raises(ValueError):
            SVIPowerLaw(**bad)
    with pytest.raises(ValueError):
        ErosionParams(law="svi").make_law()
    with pytest.raises(ValueError):
        ErosionParams(law="nope").make_law()
    p = ErosionParams(law="gauss", alpha=0.25)
    assert abs(p.make_law(2.0).delta - 0.25) < 1e-15


def test_n_cap_equals_the_oracle_bisection():
    """svimx_law_init closes a Newton bracket to adjacent doubles; the reference defines
    n_cap as the limit of a 200-step bisection (svierosion/flows.py:71-80), which the
    oracle restates: the two agree to the last bit, and V(n_cap) = 100 V(1/2)."""
    from oracle import flow as F
    from svierosion_amd.flows import SVIPowerLaw, flat_speed
    for k in range(4):
        for t in (0.51, 0.8, 0.93, 0.95, 0.999):
            for ratio in (100.0, 7.0, 1.0):
                a = SVIPowerLaw(2.0, k, 1, t, max_speed_ratio=ratio)
                b = F.SVILaw(2.0, k, 1, t, max_speed_ratio=ratio)
                assert a.n_cap == b.n_cap, (k, t, ratio, a.n_cap, b.n_cap)
                if ratio > 1:
                    assert abs(a._raw(a.n_cap) / flat_speed(a) / ratio - 1) < 1e-11
    assert SVIPowerLaw(2.0, 1, 0).n_cap == 1.0


def test_grid_sizing_equals_the_oracle():
    """svimx_grid_around / svimx_smooth_size (C++ host) against the oracle's restatement
    of svierosion/grid.py:18-56: identical shapes and bit-identical origins."""
    from oracle import stencils as S
    from svierosion_amd.grid import Grid, next_fast_size
    assert [next_fast_size(n) for n in range(1, 2000)] == [S.next_fast_size(n) for n in range(1, 2000)]
    rng = np.random.default_rng(0)
    for _ in range(500):
        lo = rng.normal(size=3) * 50
        hi = lo + rng.random(3) * 100
        h, pad = rng.uniform(0.01, 2), rng.uniform(0, 10)
        g = Grid.around(lo, hi, h, pad)
        o, _, shp = S.grid_around(lo, hi, h, pad)
        assert tuple(shp) == g.shape and np.array_equal(np.asarray(o), np.asarray(g.origin))
    g = Grid.around([0, 0, 0], [1, 1, 1], 0.1, 0.0, fast=False)
    assert g.shape == (11, 11, 11)
    with pytest.raises(ValueError):
        Grid.around([0, 0, 0], [1, 1, 1], -0.1, 0.0)
    import torch
    idx = torch.tensor([0, 5, 11 * 11 * 3 + 11 * 4 + 7])
    assert g.node_coords(idx).tolist()[2] == pytest.approx([0.3, 0.4, 0.7])


def test_marching_tetrahedra_closed_and_accurate():
    from oracle.surface import marching_tetrahedra
    from svierosion_amd.meshio import TriMesh
    h = 0.04
    ax = np.arange(-1.3, 1.3 + 1e-9, h)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    phi = (np.sqrt(X ** 2 + (1.2 * Y) ** 2 + Z ** 2) - 1.0).astype(np.float32)
    v, f = marching_tetrahedra(phi)
    m = TriMesh(-1.3 + h * v, f)
    assert m.is_watertight() and m.is_consistently_oriented()
    assert abs(m.signed_volume() / (4 / 3 * np.pi / 1.2) - 1) < 2e-3


def test_workloads_are_seeded_and_prefix_stable():
    from svierosion_amd.workloads import _job
    a = _job(4, 7, 96, 320, 0.9)
    b = _job(4, 7, 96, 320, 0.9)
    assert a[1] == b[1]
    assert np.array_equal(a[0].vertices, b[0].vertices)
    res = [_job(4, i, 96, 320, 0.9)[1].resolution for i in range(20)]
    assert min(res) >= 96 and max(res) <= 320


def test_job_device_bytes_scales_with_grid():
    from svierosion_amd.api import ErosionParams
    from svierosion_amd.batch import job_device_bytes
    from svierosion_amd.shapes import icosphere
    m = icosphere(3, 1.0)
    small = job_device_bytes(m, ErosionParams(law="svi", radius=0.1, resolution=64))
    big = job_device_bytes(m, ErosionParams(law="svi", radius=0.1, resolution=128))
    assert 4.0 < big / small < 9.0  # ~ the node count ratio (160^3 / 90^3 = 5.6)
    assert 15 < small / 90 ** 3 < 80  # tens of bytes per node, SDF transients included


def test_result_files_roundtrip():
    """The result writers produce the reference's file set (svierosion/cli.py:91-128):
    final and snapshot PLY with the quality scalar, history CSV, JSON record."""
    import json
    from pathlib import Path
    from svierosion_amd.api import ErosionResult
    from svierosion_amd.evolve import StepRecord
    from svierosion_amd.grid import Grid
    from svierosion_amd.plyio import read_ply
    from svierosion_amd.results import read_history, write_result
    from svierosion_amd.shapes import cube
    m = cube(1.0)
    m.vertex_data["depth"] = np.linspace(0.0, 1.0, len(m.vertices)).astype(np.float32)
    hist = [StepRecord(step=i + 1, time=0.1 * (i + 1), dt=0.1, volume=1.0 - 0.01 * i, fraction=1.0 - 0.01 * i,
                       vmax=2.0, vmean=1.0, qmin=0.4, qmax=0.6, capped=i, reinit=(i == 1), wall=0.0)
            for i in range(3)]
    res = ErosionResult(mesh=m, snapshots=[(0.97, 0.2, m)], history=hist, grid=Grid((0.0, 0.0, 0.0), 0.1, (8, 8, 8)),
                        stop_reason="target_volume", time=0.3, fraction=0.98, initial_volume_mesh=1.0,
                        timings={"sdf": 0.1, "evolve": 0.2, "mesh": 0.3}, params={"law": "svi", "k": 1})
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "run.ply"
        written = write_result(res, out, quality="depth")
        assert [p.name for p in written] == ["run_v0.970.ply", "run.ply"]
        back = read_ply(out)
        assert np.allclose(back.vertices, m.vertices) and np.array_equal(back.faces, m.faces)
        assert np.allclose(back.vertex_data["quality"], m.vertex_data["depth"])
        assert read_history(Path(d) / "run_history.csv") == hist
        meta = json.loads((Path(d) / "run.json").read_text())
        assert meta["stop_reason"] == "target_volume" and meta["grid"]["shape"] == [8, 8, 8]
        assert meta["snapshots"] == [{"fraction": 0.97, "time": 0.2}]
        assert abs(meta["final_mesh_volume"] - 1.0) < 1e-9


def _np_edge_topology(faces, nv):
    """Edge counts by sorting (a NumPy restatement of svierosion/meshio.py:88-105)."""
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    lo, hi = e.min(1), e.max(1)
    uniq, inv, counts = np.unique(lo * max(nv, 1) + hi, return_inverse=True, return_counts=True)
    bal = np.bincount(inv.reshape(-1), weights=np.where(e[:, 0] < e[:, 1], 1, -1), minlength=len(uniq))
    return {"edges": len(uniq), "boundary_edges": int((counts == 1).sum()),
            "nonmanifold_edges": int((counts > 2).sum()), "misoriented_edges": int(((counts == 2) & (bal != 0)).sum())}


def test_mesh_measures_match_numpy():
    """The library's host loops for the mesh volume and the edge counts agree
    with NumPy restatements on closed, flipped, holed and non-manifold meshes;
    the volume also against closed forms (unit cube 1, icosphere -> 4/3 pi)."""
    from svierosion_amd import shapes
    from svierosion_amd.meshio import TriMesh
    cube = shapes.cube(1.0, center=(3.0, -2.0, 5.0))
    assert abs(cube.signed_volume() - 1.0) < 1e-12
    assert abs(cube.flipped().signed_volume() + 1.0) < 1e-12
    ico = shapes.icosphere(5, 1.0)
    assert abs(ico.volume() - 4.0 / 3.0 * np.pi) < 5e-3  # inscribed polyhedron
    torus = shapes.torus()
    rng = np.random.default_rng(3)
    cases = [cube, ico, torus]
    f = ico.faces.copy()
    idx = rng.choice(len(f), 7, replace=False)
    f[idx] = f[idx][:, ::-1]  # flipped faces
    cases.append(TriMesh(ico.vertices, f))
    cases.append(TriMesh(ico.vertices, ico.faces[5:]))  # holes
    extra = np.array([[ico.faces[0, 0], ico.faces[0, 1], 0]])  # a third face on an edge
    cases.append(TriMesh(ico.vertices, np.concatenate([ico.faces, extra])))
    for m in cases:
        t = m.triangles() - m.vertices.mean(0)
        want = float(np.einsum("ij,ij->i", t[:, 0], np.cross(t[:, 1], t[:, 2])).sum() / 6.0)
        assert abs(m.signed_volume() - want) <= 1e-12 * max(1.0, abs(want))
        assert m.edge_topology() == _np_edge_topology(m.faces, len(m.vertices))
    assert cases[0].is_watertight() and cases[0].is_consistently_oriented()
    assert not cases[3].is_consistently_oriented()
    assert cases[4].edge_topology()["boundary_edges"] > 0
    assert cases[5].edge_topology()["nonmanifold_edges"] == 1
    bad = TriMesh(ico.vertices, ico.faces)
    bad.faces = bad.faces.copy()
    bad.faces[0, 0] = len(ico.vertices)  # out of range: refused, no crash
    with pytest.raises(ValueError, match="vertex"):
        bad.signed_volume()
    with pytest.raises(ValueError, match="vertex"):
        bad.edge_topology()


def test_sdf_prepare_native_matches_numpy():
    """The library's SDF preparation equals the NumPy path for meshes without
    long edges: index coordinates, frames and boxes exactly; vertex ids as the
    same grouping (bitwise-equal vertices of a triangle soup share one), and
    the same star triangles per vertex in corner order.  A mesh with a long
    edge is left to the NumPy splitting path."""
    from svierosion_amd import shapes, sdf
    from svierosion_amd.api import ErosionParams, make_grid
    from svierosion_amd.meshio import TriMesh
    ico = shapes.icosphere(4, 1.0)
    soup = TriMesh(ico.vertices[ico.faces].reshape(-1, 3), np.arange(3 * len(ico.faces)).reshape(-1, 3))
    for mesh in (ico, soup):
        g = make_grid(mesh, ErosionParams(law="mean", resolution=48))
        nat = sdf._prepare_native(mesh, g, 2.0)
        assert nat is not None
        # the NumPy path, with the native one bypassed
        saved = sdf._prepare_native
        sdf._prepare_native = lambda *a: None
        try:
            old = sdf._prepare(mesh, g, 2.0)
        finally:
            sdf._prepare_native = saved
        for a, b in zip(nat[:5], old[:5]):
            assert np.array_equal(np.asarray(a, np.float64), np.asarray(b, np.float64))
        vn, vo = nat[5].reshape(-1), old[5].reshape(-1)
        # same grouping: a bijection between the id sets
        pairs = set(zip(vn.tolist(), vo.tolist()))
        assert len(pairs) == len(set(vn.tolist())) == len(set(vo.tolist()))
        assert len(set(vo.tolist())) == len(ico.vertices)  # the soup's corners regroup into the sphere's vertices
        m = dict(pairs)
        for idn in range(len(nat[7])):
            a = nat[6][nat[8][idn]: nat[8][idn] + nat[7][idn]]
            ido = m[idn]
            b = old[6][old[8][ido]: old[8][ido] + old[7][ido]]
            assert np.array_equal(a, b)
    tri = shapes.cube(1.0)
    g = make_grid(tri, ErosionParams(law="mean", resolution=48))
    assert sdf._prepare_native(tri, g, 2.0) is None  # 48-cell edges: split on the NumPy path


def test_fifo_replay_follows_the_single_server_queue(monkeypatch):
    """serve_sequential(wait_for_arrivals=False): requests run back to back and the FIFO
    queue is replayed from their service times -- start = max(arrival, previous finish),
    finish = start + service (no GPU: erode is replaced by a timed stand-in)."""
    import time as _time
    from types import SimpleNamespace
    from svierosion_amd import batch
    service = {0: 0.03, 1: 0.01, 2: 0.02, 3: 0.01}

    def fake_erode(mesh, params, keep_state=False, callback=None):
        _time.sleep(service[params])
        return SimpleNamespace(history=[0] * (params + 1))

    monkeypatch.setattr(batch, "erode", fake_erode)
    arrivals = [0.0, 0.005, 0.5, 0.51]
    reqs = [batch.Request(None, i, a) for i, a in enumerate(arrivals)]
    rep = batch.serve_sequential(reqs, wait_for_arrivals=False)
    c = rep.completions
    assert [x.index for x in c] == [0, 1, 2, 3]
    assert c[0].admitted == 0.0 and c[1].admitted == pytest.approx(c[0].finished)
    assert c[2].admitted == 0.5 and c[3].admitted == pytest.approx(c[2].finished)
    for x in c:
        assert x.finished - x.admitted == pytest.approx(service[x.index], abs=0.02)
        assert x.steps == x.index + 1
    assert rep.makespan == pytest.approx(c[3].finished)
