"""Node grids: the ``Grid`` type, ``get_device`` and ``next_fast_size`` of
svierosion/grid.py.

A grid is an axis-aligned lattice of nx*ny*nz nodes at origin + spacing*(i, j, k),
stored in C order over (x, y, z); node (i, j, k) also stands for the cube of side
``spacing`` centred on it.  The sizing rule that places a mesh's bounding box in
a grid (``Grid.around``) runs in the library (``svimx_grid_around``, FP64 in the
reference's operation order) because its output -- shape and origin -- fixes
every node position that parity compares.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi


def get_device(device=None) -> torch.device:
    """The ROCm device for ``None`` or ``"auto"`` when one is present (the CPU
    otherwise, which the compute entry points then refuse); any other value is
    taken as a torch device name."""
    if device in (None, "auto"):
        device = "cuda" if torch.cuda.is_available() else "cpu"
    return torch.device(device)


def next_fast_size(n: int) -> int:
    """Smallest size >= n with no prime factor above 7."""
    out = C.c_int64()
    _abi.check(_abi.lib.svimx_smooth_size(int(n), C.byref(out)))
    return int(out.value)


@dataclass(frozen=True)
class Grid:
    """origin: world position of node (0, 0, 0); spacing: node distance h;
    shape: (nx, ny, nz)."""

    origin: tuple
    spacing: float
    shape: tuple

    @classmethod
    def around(cls, lo, hi, spacing: float, padding: float, fast: bool = True) -> "Grid":
        """The grid holding the box [lo, hi] at its centre with at least ``padding``
        of free space on every side (dimensions 7-smooth when ``fast``)."""
        lo3 = (C.c_double * 3)(*np.asarray(lo, dtype=np.float64).reshape(3))
        hi3 = (C.c_double * 3)(*np.asarray(hi, dtype=np.float64).reshape(3))
        shape = (C.c_int64 * 3)()
        origin = (C.c_double * 3)()
        _abi.check(_abi.lib.svimx_grid_around(lo3, hi3, float(spacing), float(padding), int(bool(fast)), shape,
                                              origin))
        return cls(tuple(origin), float(spacing), tuple(int(n) for n in shape))

    # -- lattice facts
    @property
    def size(self) -> int:
        return int(np.prod(np.asarray(self.shape, dtype=np.int64)))

    @property
    def strides(self) -> tuple:
        """Flat-index steps of one node along x, y, z."""
        ny, nz = int(self.shape[1]), int(self.shape[2])
        return (ny * nz, nz, 1)

    def _o(self) -> np.ndarray:
        return np.asarray(self.origin, dtype=np.float64)

    # -- coordinates
    def axes(self) -> list:
        """Node coordinates along each axis."""
        return [self.origin[a] + self.spacing * np.arange(self.shape[a]) for a in range(3)]

    def to_index(self, points) -> np.ndarray:
        """World points -> continuous index coordinates (cells from node 0)."""
        return (np.asarray(points, dtype=np.float64) - self._o()) / self.spacing

    def to_world(self, index) -> np.ndarray:
        """Index coordinates -> world points."""
        return self._o() + np.asarray(index, dtype=np.float64) * self.spacing

    def node_coords(self, flat_index: torch.Tensor, dtype=torch.float64) -> torch.Tensor:
        """World positions (m, 3) of flat node indices."""
        o = torch.as_tensor(self._o(), dtype=dtype, device=flat_index.device)
        return o + unravel(flat_index, self.shape).to(dtype) * self.spacing

    def bounds(self) -> np.ndarray:
        """[[first node], [last node]] in world coordinates."""
        last = np.asarray(self.shape, dtype=np.float64) - 1.0
        return np.stack([self._o(), self.to_world(last)])

    def describe(self) -> str:
        nx, ny, nz = self.shape
        return f"{nx}x{ny}x{nz} grid ({self.size / 1e6:.2f}M nodes, h = {self.spacing:.6g})"


def unravel(flat: torch.Tensor, shape) -> torch.Tensor:
    """Flat C-order node indices -> (m, 3) integer (i, j, k)."""
    ny, nz = int(shape[1]), int(shape[2])
    flat = flat.to(torch.int64)
    rest = torch.div(flat, nz, rounding_mode="floor")
    k = flat - rest * nz
    i = torch.div(rest, ny, rounding_mode="floor")
    j = rest - i * ny
    return torch.stack([i, j, k], dim=1)
