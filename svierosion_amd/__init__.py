"""svierosion_amd: MI355X-native narrow-band SVI erosion.

Drop-in for the erosion path of jwcalder/SVIErosion: the same
``erode(mesh, ErosionParams)`` / ``TriMesh`` API (svierosion/__init__.py:20-34)
over hand-written HIP kernels for gfx950 (``libsvimx.so``, C ABI in
``include/svimx.h``).  Names are resolved lazily so that
``python -m svierosion_amd.build`` works before the library exists.
"""

from __future__ import annotations

import importlib

__version__ = "0.1.0"

_EXPORTS = {
    "ErosionParams": "api", "ErosionResult": "api", "erode": "api", "point_mesh_distance": "api",
    "make_grid": "api", "volume_fraction_report": "api",
    "ErosionSimulation": "evolve", "StepRecord": "evolve", "RunResult": "evolve",
    "SVIPowerLaw": "flows", "MeanCurvatureLaw": "flows", "GaussCurvatureLaw": "flows", "make_law": "flows",
    "Grid": "grid", "get_device": "grid",
    "LevelSet": "levelset",
    "TriMesh": "meshio",
    "mesh_to_sdf": "sdf",
    "extract_surface": "surface",
    "BallConvolution": "svi",
    "read_ply": "plyio", "write_ply": "plyio",
}

__all__ = sorted(_EXPORTS)


def __getattr__(name):
    mod = _EXPORTS.get(name)
    if mod is None:
        raise AttributeError(name)
    return getattr(importlib.import_module(f".{mod}", __name__), name)
