"""Uniform Cartesian grids (API of svierosion/grid.py).

Grid sizing must match the reference exactly -- the node positions, shape and
origin define every field compared in parity tests -- so ``next_fast_size``
and ``Grid.around`` implement the same arithmetic as svierosion/grid.py:18-56.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch


def get_device(device=None) -> torch.device:
    """``None``/``"auto"`` -> the ROCm device when present, else the CPU
    (svierosion/grid.py:11-15).  Compute entry points reject the CPU."""
    if device is None or device == "auto":
        return torch.device("cuda" if torch.cuda.is_available() else "cpu")
    return torch.device(device)


def next_fast_size(n: int) -> int:
    """Smallest m >= n whose prime factors are all in {2, 3, 5, 7}."""
    m = max(int(n), 1)
    while True:
        r = m
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 1


@dataclass(frozen=True)
class Grid:
    """Nodes ``origin + spacing * (i, j, k)``, 0 <= (i, j, k) < shape, C order (x, y, z)."""

    origin: tuple
    spacing: float
    shape: tuple

    @classmethod
    def around(cls, lo, hi, spacing: float, padding: float, fast: bool = True) -> "Grid":
        """Grid holding [lo, hi] centred, with >= ``padding`` on each side."""
        lo = np.asarray(lo, dtype=np.float64)
        hi = np.asarray(hi, dtype=np.float64)
        need = np.ceil((hi - lo + 2 * padding) / spacing).astype(int) + 1
        shape = tuple(int(next_fast_size(n)) if fast else int(n) for n in need)
        centre = 0.5 * (lo + hi)
        origin = centre - 0.5 * (np.asarray(shape, dtype=np.float64) - 1) * spacing
        return cls(tuple(float(o) for o in origin), float(spacing), shape)

    @property
    def size(self) -> int:
        nx, ny, nz = self.shape
        return int(nx) * int(ny) * int(nz)

    @property
    def strides(self) -> tuple:
        _, ny, nz = self.shape
        return (ny * nz, nz, 1)

    def axes(self) -> list:
        return [self.origin[a] + self.spacing * np.arange(self.shape[a]) for a in range(3)]

    def to_index(self, points) -> np.ndarray:
        return (np.asarray(points, dtype=np.float64) - np.asarray(self.origin)) / self.spacing

    def to_world(self, index) -> np.ndarray:
        return np.asarray(self.origin) + self.spacing * np.asarray(index, dtype=np.float64)

    def node_coords(self, flat_index: torch.Tensor, dtype=torch.float64) -> torch.Tensor:
        ijk = unravel(flat_index, self.shape).to(dtype)
        o = torch.tensor(self.origin, dtype=dtype, device=flat_index.device)
        return o + self.spacing * ijk

    def bounds(self) -> np.ndarray:
        o = np.asarray(self.origin)
        return np.stack([o, o + self.spacing * (np.asarray(self.shape) - 1)])

    def describe(self) -> str:
        return (f"grid {self.shape[0]}x{self.shape[1]}x{self.shape[2]} "
                f"({self.size / 1e6:.1f}M nodes), spacing {self.spacing:.6g}")


def unravel(flat: torch.Tensor, shape) -> torch.Tensor:
    """Flat C-order indices -> (m, 3) integer indices."""
    _, ny, nz = shape
    flat = flat.to(torch.int64)
    return torch.stack([flat // (ny * nz), (flat // nz) % ny, flat % nz], dim=1)
