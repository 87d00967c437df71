"""Narrow-band level set on the device (API of svierosion/levelset.py).

Device layout (per job, in HBM):

* three dense float32 copies of phi (``bufs``); one holds the current field,
  the other two are the TVD-RK3 stage outputs.  Outside the band every copy
  holds the same clamped value +-width*h, so any copy can serve as a stencil
  input; ``rebuild_band`` re-synchronises them.
* ``occupancy`` -- the occupancy f (1 inside, 0 outside, the tangent-plane
  cell fraction at band nodes), the input of the N_r summation, stored in the
  zero-padded layout of ``svimx_occupancy_elems``.  Only band entries change
  between rebuilds.
* the band -- a sorted int32 list of flat indices (capacity = grid size).
* a workspace with the reduction partials and the result slots.

All numerics run in libsvimx; this class sequences the calls.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi
from .grid import Grid

MARGIN = 4  # cells between any band node and the grid boundary


def _idx32(idx: torch.Tensor, device) -> torch.Tensor:
    return idx.to(device=device, dtype=torch.int32).contiguous()


class LevelSet:
    """A level set function with an explicitly stored narrow band."""

    def __init__(self, phi: torch.Tensor, grid: Grid, width: float = 6.0):
        if tuple(phi.shape) != tuple(grid.shape):
            raise ValueError("phi shape does not match grid")
        self.device = _abi.require_device(phi.device if phi.is_cuda else "cuda")
        self.grid = grid
        self.h = float(grid.spacing)
        self.width = float(width)
        self.dims = tuple(int(s) for s in grid.shape)
        self.strides = tuple(int(s) for s in grid.strides)
        n = grid.size
        if n >= 2 ** 31:
            raise ValueError("grid has more than 2^31 nodes")
        first = phi.to(self.device, torch.float32).contiguous().reshape(-1).clone()
        self.bufs = [first, torch.empty_like(first), torch.empty_like(first)]
        self.cur = 0
        self.occupancy = _abi.occupancy_zeros(self.dims, self.device)
        self.ws = _abi.workspace(self.dims, self.device)
        self._band = torch.empty(n, dtype=torch.int32, device=self.device)
        self.nb = 0
        self.inside_outside = 0
        self._reinit_bufs = None
        self._reinit_work = None
        self.rebuild_band()

    # ------------------------------------------------------------ state
    @property
    def flat(self) -> torch.Tensor:
        return self.bufs[self.cur]

    @property
    def phi(self) -> torch.Tensor:
        return self.flat.view(self.dims)

    @property
    def band(self) -> torch.Tensor:
        """Flat indices of the band nodes, sorted, as int64 (the reference's
        ``LevelSet.band`` type, svierosion/levelset.py:54-86): a copy of the
        resident int32 list that the kernels read (``band32``)."""
        return self._band[: self.nb].long()

    @property
    def band32(self) -> torch.Tensor:
        """The resident int32 band list itself (a view, no copy)."""
        return self._band[: self.nb]

    def results(self) -> np.ndarray:
        """Host copy of the workspace result slots (synchronises the stream)."""
        return _abi.results_view(self.ws).cpu().numpy()

    # ------------------------------------------------------------ band
    def rebuild_band(self) -> None:
        """Band = nodes with |phi| < width*h; clamp all other values (svierosion/levelset.py:70-86)."""
        others = [b for i, b in enumerate(self.bufs) if i != self.cur]
        info = (C.c_int64 * 9)()
        rc = _abi.lib.svimx_band_rebuild(_abi.stream(), _abi.ptr(self.flat), _abi.ptr(others[0]),
                                         _abi.ptr(others[1]), _abi.ptr(self.occupancy), *self.dims, self.h,
                                         self.width, MARGIN, _abi.ptr(self._band), self._band.numel(),
                                         _abi.ptr(self.ws), self.ws.numel(), info)
        self.nb = int(info[0])
        self.inside_outside = int(info[1])
        _abi.check(rc, "band rebuild")

    def fractions(self, plan=None, cut_hi: float = 5.5) -> None:
        """Cell fractions of the current band into ``occupancy`` (+ volume and interface results)."""
        _abi.check(_abi.lib.svimx_fractions(_abi.stream(), _abi.ptr(self.flat), _abi.ptr(self._band), self.nb,
                                            *self.dims, self.h, _abi.ptr(self.occupancy), plan, float(cut_hi),
                                            _abi.ptr(self.ws), self.ws.numel()), "fractions")

    def volume_from_results(self, res) -> float:
        return (self.inside_outside + float(res[_abi.R_VOLUME_SUM])) * self.h ** 3

    def interface_nodes(self, idx: torch.Tensor | None = None) -> torch.Tensor:
        """Band nodes having a 6-neighbour of the opposite (or zero) sign."""
        idx = self.band32 if idx is None else _idx32(idx, self.device)
        out = torch.empty(idx.numel(), dtype=torch.uint8, device=self.device)
        _abi.check(_abi.lib.svimx_op_interface_mask(_abi.stream(), _abi.ptr(self.flat), _abi.ptr(idx), idx.numel(),
                                                    *self.dims, _abi.ptr(out)))
        return idx[out.bool()]

    def has_interface(self) -> bool:
        self.fractions()
        r = self.results()
        return bool(r[_abi.R_HAS_NEG] > 0) and bool(r[_abi.R_HAS_POS] > 0)

    # ---------------------------------------------------- derivatives
    def central_gradient(self, idx: torch.Tensor) -> torch.Tensor:
        f = self.flat
        idx = idx.to(self.device, torch.int64)
        return torch.stack([0.5 * (f[idx + s] - f[idx - s]) for s in self.strides], 1)

    def godunov_norm_inward(self, idx: torch.Tensor) -> torch.Tensor:
        idx = _idx32(idx, self.device)
        out = torch.empty(idx.numel(), dtype=torch.float32, device=self.device)
        _abi.check(_abi.lib.svimx_op_godunov(_abi.stream(), _abi.ptr(self.flat), _abi.ptr(idx), idx.numel(),
                                             *self.dims, 1.0 / self.h, _abi.ptr(out)))
        return out

    def principal_curvatures(self, idx: torch.Tensor):
        idx = _idx32(idx, self.device)
        k1 = torch.empty(idx.numel(), dtype=torch.float32, device=self.device)
        k2 = torch.empty_like(k1)
        _abi.check(_abi.lib.svimx_op_principal_curvatures(_abi.stream(), _abi.ptr(self.flat), _abi.ptr(idx),
                                                          idx.numel(), *self.dims, self.h, _abi.ptr(k1),
                                                          _abi.ptr(k2)))
        return k1, k2

    # ------------------------------------------------- volume fractions
    def cell_fractions(self, idx: torch.Tensor | None = None) -> torch.Tensor:
        idx = self.band32 if idx is None else _idx32(idx, self.device)
        out = torch.empty(idx.numel(), dtype=torch.float32, device=self.device)
        _abi.check(_abi.lib.svimx_op_cell_fractions(_abi.stream(), _abi.ptr(self.flat), _abi.ptr(idx),
                                                    idx.numel(), *self.dims, _abi.ptr(out)))
        return out

    def dense_fractions(self) -> torch.Tensor:
        """Dense occupancy (1 inside, 0 outside, cell fractions in the band)."""
        self.fractions()
        return _abi.occupancy_view(self.occupancy, self.dims).clone()

    def volume(self) -> float:
        self.fractions()
        return self.volume_from_results(self.results())

    # ----------------------------------------------------- projection
    def closest_points(self, idx: torch.Tensor, newton_steps: int = 1) -> torch.Tensor:
        """Capped Newton projection onto the zero set, index coordinates (m, 3): the
        first step with the central gradient at the node, each further step with the
        tricubic value and gradient at the current point (svierosion/levelset.py:129-151)."""
        idx = _idx32(idx, self.device)
        out = torch.empty((idx.numel(), 3), dtype=torch.float32, device=self.device)
        _abi.check(_abi.lib.svimx_closest_points_newton(_abi.stream(), _abi.ptr(self.flat), _abi.ptr(idx),
                                                        idx.numel(), *self.dims, self.h, int(newton_steps),
                                                        _abi.ptr(out)), "closest points")
        return out

    def clamp_to_interior(self, p: torch.Tensor) -> torch.Tensor:
        hi = torch.tensor([n - 3.001 for n in self.dims], device=p.device, dtype=p.dtype)
        return torch.minimum(p.clamp_min(1.0), hi)

    # --------------------------------------------------- reinitialisation
    def reinitialize(self, iterations: int = 12, cfl: float = 0.45, dilation: int = 3, gate=None) -> bool:
        """Min (2010) PDE reinitialisation with subcell fix, then band rebuild
        (svierosion/levelset.py:154-183).  ``gate`` (an ``_abi.ReinitGate``): run
        only if the device-side trigger over the workspace's step results opens
        (svimx_reinit_gated); returns whether it ran."""
        n = self.grid.size
        self.reserve_reinit()
        ma, mb = self._reinit_bufs
        others = [b for i, b in enumerate(self.bufs) if i != self.cur]
        info = (C.c_int64 * 11)()
        cap = min(n, 2 * self.nb + 4096)
        for _ in range(2):
            widx, winfo, geo = self._work_buffers(cap)
            args = (_abi.stream(), _abi.ptr(self.flat), _abi.ptr(others[0]), _abi.ptr(others[1]),
                    _abi.ptr(self.occupancy), _abi.ptr(self._band), self.nb, self.inside_outside, *self.dims, self.h,
                    self.width, int(iterations), float(cfl), int(dilation), MARGIN, _abi.ptr(ma), _abi.ptr(mb),
                    _abi.ptr(widx), _abi.ptr(winfo), _abi.ptr(geo), cap, _abi.ptr(self.ws), self.ws.numel())
            if gate is None:
                rc = _abi.lib.svimx_reinit(*args, info)
            else:
                rc = _abi.lib.svimx_reinit_gated(*args, C.byref(gate), info)
            if rc != _abi.ECAPACITY:
                break
            cap = int(info[9])  # the work set did not fit: retry with its exact size
        ran = gate is None or int(info[10]) == 1
        if ran and rc in (_abi.OK, _abi.EMARGIN):
            self.nb = int(info[0])
            self.inside_outside = int(info[1])
            self.last_work_nodes = int(info[9])
        _abi.check(rc, "reinitialisation")
        return ran

    def reserve_reinit(self) -> None:
        """Allocate the reinitialisation's masks and work buffers now (for the
        current band, with the usual headroom) rather than at the first
        reinitialisation."""
        n = self.grid.size
        if self._reinit_bufs is None:
            self._reinit_bufs = (torch.empty(n, dtype=torch.uint8, device=self.device),
                                 torch.empty(n, dtype=torch.uint8, device=self.device))
        self._work_buffers(min(n, 2 * self.nb + 4096))

    def _work_buffers(self, cap: int):
        """Work list, per-node flags and geometry for a work set of ``cap`` nodes,
        kept between reinitialisations and grown with 25% headroom, so the band's
        drift does not reallocate them every time."""
        have = self._reinit_work[0].numel() if self._reinit_work is not None else -1
        if have < cap:
            size = min(self.grid.size, cap + cap // 4)
            self._reinit_work = None  # release the old buffers before the new ones
            self._reinit_work = (torch.empty(size, dtype=torch.int32, device=self.device),
                                 torch.empty(size, dtype=torch.int32, device=self.device),
                                 torch.empty(8 * size, dtype=torch.float32, device=self.device))
        return self._reinit_work


# ------------------------------------------------------------- interpolation
def _interp(fn, flat: torch.Tensor, shape, p: torch.Tensor) -> torch.Tensor:
    f = flat.contiguous().reshape(-1)
    p = p.to(f.device, torch.float32).contiguous()
    out = torch.empty(p.shape[0], dtype=torch.float32, device=f.device)
    _abi.check(fn(_abi.stream(), _abi.ptr(f), *map(int, shape), _abi.ptr(p), p.shape[0], _abi.ptr(out)))
    return out


def tricubic(flat: torch.Tensor, shape, p: torch.Tensor) -> torch.Tensor:
    """Tensor-product cubic Lagrange interpolation at index coordinates p (m, 3)."""
    return _interp(_abi.lib.svimx_op_tricubic, flat, shape, p)


def trilinear(flat: torch.Tensor, shape, p: torch.Tensor) -> torch.Tensor:
    """Trilinear interpolation at index coordinates p (m, 3)."""
    return _interp(_abi.lib.svimx_op_trilinear, flat, shape, p)
