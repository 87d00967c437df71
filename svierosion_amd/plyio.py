"""PLY reading and writing (the on-disk format of svierosion/meshio.py:135-361).

Reads ascii, binary_little_endian and binary_big_endian files with arbitrary
scalar vertex properties (x, y, z plus extra per-vertex scalars, returned in
``vertex_data``) and polygonal faces given as a list property (fan
triangulated).  Writes binary little-endian (float or double coordinates) or
ascii, with optional float32 vertex scalars.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from .meshio import TriMesh

_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
          "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
          "float": "f4", "float32": "f4", "double": "f8", "float64": "f8"}


def _header(fh):
    if fh.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt, elements = None, []
    while True:
        line = fh.readline()
        if not line:
            raise ValueError("truncated PLY header")
        tok = line.decode("ascii", errors="replace").split()
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append({"name": tok[1], "count": int(tok[2]), "props": []})
        elif tok[0] == "property":
            if tok[1] == "list":
                elements[-1]["props"].append((tok[4], _TYPES[tok[3]], _TYPES[tok[2]]))
            else:
                elements[-1]["props"].append((tok[2], _TYPES[tok[1]], None))
        elif tok[0] == "end_header":
            break
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise ValueError(f"unsupported PLY format {fmt!r}")
    return fmt, elements


def _fan(polys):
    tris = []
    for p in polys:
        for q in range(1, len(p) - 1):
            tris.append((p[0], p[q], p[q + 1]))
    return np.asarray(tris, dtype=np.int64).reshape(-1, 3)


def _read_binary(fh, elements, order):
    data = {}
    for el in elements:
        props = el["props"]
        if all(p[2] is None for p in props):
            dt = np.dtype([(name, order + t) for name, t, _ in props])
            arr = np.frombuffer(fh.read(dt.itemsize * el["count"]), dtype=dt, count=el["count"])
            data[el["name"]] = {name: arr[name] for name, _, _ in props}
            continue
        cols = {name: [] for name, _, _ in props}
        for _ in range(el["count"]):
            for name, t, ct in props:
                if ct is None:
                    cols[name].append(np.frombuffer(fh.read(np.dtype(t).itemsize), order + t)[0])
                else:
                    n = int(np.frombuffer(fh.read(np.dtype(ct).itemsize), order + ct)[0])
                    cols[name].append(np.frombuffer(fh.read(np.dtype(t).itemsize * n), order + t, count=n))
        data[el["name"]] = cols
    return data


def _read_ascii(fh, elements):
    words = fh.read().decode("ascii", errors="replace").split()
    pos = 0
    data = {}
    for el in elements:
        cols = {name: [] for name, _, _ in el["props"]}
        for _ in range(el["count"]):
            for name, t, ct in el["props"]:
                if ct is None:
                    cols[name].append(float(words[pos]))
                    pos += 1
                else:
                    n = int(float(words[pos]))
                    cols[name].append([int(float(w)) for w in words[pos + 1: pos + 1 + n]])
                    pos += 1 + n
        data[el["name"]] = cols
    return data


def read_ply(path, vertex_properties=True) -> TriMesh:
    """Triangle mesh from a PLY file (polygons fan-triangulated).

    ``vertex_properties``: True keeps every extra scalar vertex property in
    ``vertex_data`` with its on-disk type, False none; a collection of names
    keeps only those."""
    with open(path, "rb") as fh:
        fmt, elements = _header(fh)
        types = {el["name"]: {name: t for name, t, _ in el["props"]} for el in elements}
        if fmt == "ascii":
            data = _read_ascii(fh, elements)
        else:
            data = _read_binary(fh, elements, "<" if fmt == "binary_little_endian" else ">")
    v = data.get("vertex")
    if v is None:
        raise ValueError("PLY file has no vertex element")
    verts = np.stack([np.asarray(v[c], dtype=np.float64) for c in ("x", "y", "z")], 1)
    faces = np.zeros((0, 3), dtype=np.int64)
    f = data.get("face")
    if f is not None:
        key = next((k for k in ("vertex_indices", "vertex_index") if k in f), None)
        if key is None:
            raise ValueError("PLY face element lists no vertex_indices / vertex_index property")
        faces = _fan([np.asarray(p, dtype=np.int64) for p in f[key]])
    if vertex_properties is True:
        keep = [k for k in v if k not in ("x", "y", "z")]
    elif not vertex_properties:
        keep = []
    else:
        keep = [k for k in v if k in set(vertex_properties) and k not in ("x", "y", "z")]
    # on-disk scalar type in native byte order
    extra = {k: np.asarray(v[k]).astype(np.dtype(types["vertex"][k]).newbyteorder("="), copy=True) for k in keep}
    return TriMesh(verts, faces, extra)


def write_ply(path, mesh: TriMesh, binary: bool = True, precision: str = "float", vertex_data=None,
              comments=None) -> None:
    """Write ``mesh`` (plus ``mesh.vertex_data`` and ``vertex_data`` as float32
    vertex scalars) as PLY, binary little-endian or ascii, creating missing parent
    directories; each entry of ``comments`` becomes a ``comment`` header line."""
    extra = dict(mesh.vertex_data)
    if vertex_data:
        extra.update(vertex_data)
    n, m = len(mesh.vertices), len(mesh.faces)
    for k, val in extra.items():
        if np.shape(val) != (n,):
            raise ValueError(f"vertex property {k!r} has shape {np.shape(val)}, expected ({n},)")
    ctype = "float" if precision == "float" else "double"
    vt = "<f4" if precision == "float" else "<f8"
    head = ["ply", f"format {'binary_little_endian' if binary else 'ascii'} 1.0"]
    head += ["comment " + str(c).replace("\r", " ").replace("\n", " ") for c in (comments or [])]
    head += [f"element vertex {n}", f"property {ctype} x", f"property {ctype} y", f"property {ctype} z"]
    head += [f"property float {k}" for k in extra]
    head += [f"element face {m}", "property list uchar int vertex_indices", "end_header"]
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    with open(Path(path), "wb") as fh:
        fh.write(("\n".join(head) + "\n").encode("ascii"))
        if binary:
            dt = np.dtype([("x", vt), ("y", vt), ("z", vt)] + [(k, "<f4") for k in extra])
            rec = np.empty(n, dtype=dt)
            for a, c in enumerate("xyz"):
                rec[c] = mesh.vertices[:, a]
            for k, val in extra.items():
                rec[k] = np.asarray(val, dtype=np.float32)
            fh.write(rec.tobytes())
            fr = np.empty(m, dtype=np.dtype([("n", "u1"), ("v", "<i4", (3,))]))
            fr["n"] = 3
            fr["v"] = mesh.faces
            fh.write(fr.tobytes())
        else:
            cols = [mesh.vertices] + [np.asarray(val, dtype=np.float64)[:, None] for val in extra.values()]
            vals = np.concatenate(cols, 1)
            fmt = "%.17g" if precision != "float" else "%.9g"
            lines = [" ".join(fmt % x for x in row) for row in vals]
            lines += [f"3 {a} {b} {c}" for a, b, c in mesh.faces]
            fh.write(("\n".join(lines) + "\n").encode("ascii"))
