"""Spherical volume invariant N_r (API of svierosion/svi.py).

For occupancy f (1 inside, tangent-plane cell fraction in the band),
|B_r(x) cap Omega| ~ h^3 sum_y w(x - y) f(y) with w the exact ball/cell
overlap fractions, and N_r = 1 - |B_r cap Omega| / |B_r|.

The weights are computed once per job on the host in FP64 (the same scheme
as svierosion/svi.py:24-55: exact classification of whole cells, analytic
z-extent times a Gauss-Legendre rule in x and y for cut cells).  The sum is
evaluated on the device by the direct tiled summation of libsvimx, only where
N_r is needed (the tricubic stencils of the closest points); see
``svimx_nr_exec``.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _abi
from .grid import Grid


def ball_cell_weights(radius: float, spacing: float, quad: int = 32) -> np.ndarray:
    """|B_r(0) cap cell(i, j, k)| / h^3 for |i|, |j|, |k| <= m, m = floor(R + 1/2) + 1."""
    R = radius / spacing
    m = int(math.floor(R + 0.5)) + 1
    off = np.arange(-m, m + 1, dtype=np.float64)
    cx, cy, cz = np.meshgrid(off, off, off, indexing="ij")
    ax, ay, az = np.abs(cx), np.abs(cy), np.abs(cz)
    near = np.sqrt(np.maximum(ax - 0.5, 0) ** 2 + np.maximum(ay - 0.5, 0) ** 2 + np.maximum(az - 0.5, 0) ** 2)
    far = np.sqrt((ax + 0.5) ** 2 + (ay + 0.5) ** 2 + (az + 0.5) ** 2)
    w = np.where(far <= R, 1.0, 0.0)
    cut = np.argwhere((near < R) & (far > R))
    nodes, gw = np.polynomial.legendre.leggauss(quad)
    nodes, gw = 0.5 * nodes, 0.5 * gw
    w2 = np.outer(gw, gw)
    for s in range(0, len(cut), 2048):
        c = cut[s:s + 2048]
        x = off[c[:, 0]][:, None, None] + nodes[None, :, None]
        y = off[c[:, 1]][:, None, None] + nodes[None, None, :]
        half = np.sqrt(np.maximum(R * R - x * x - y * y, 0.0))      # ball's z half-extent
        z = off[c[:, 2]][:, None, None]
        seg = np.clip(np.minimum(z + 0.5, half) - np.maximum(z - 0.5, -half), 0.0, None)
        w[c[:, 0], c[:, 1], c[:, 2]] = (seg * w2).sum((1, 2))
    return w


def sphere_exterior_fraction(radius_ball: float, radius_sphere: float) -> float:
    """Exact N_r on a sphere of radius R (valid for r <= 2R)."""
    r, R = radius_ball, radius_sphere
    if r > 2 * R:
        return 1.0 - (R / r) ** 3
    return 0.5 + 3.0 * r / (16.0 * R)


_TUNE_SHAPES = 11  # SVIMX_NR_TUNE_SHAPES


def tune_shape(i: int) -> tuple:
    """(run length, lines, lds) of the autotuner's shape i."""
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    _abi.check(_abi.lib.svimx_nr_tune_shape(int(i), C.byref(a), C.byref(b), C.byref(c)), "tuning shape")
    return a.value, b.value, bool(c.value)


def _shape_name(rl: int, lines: int, lds: bool) -> str:
    return f"{rl}/lds" if lds else (f"{rl}" if lines <= 1 else f"{rl}x{lines}")


class BallConvolution:
    """Ball weights of a grid as a libsvimx plan; evaluates N_r on the device.

    Attributes mirror svierosion/svi.py:58-82: ``m``, ``ball_volume``
    (4/3 pi r^3, exact) and ``quadrature_volume`` (h^3 sum w).
    """

    def __init__(self, grid: Grid, radius: float, device=None, quad: int = 32, method: str = "auto"):
        self.grid = grid
        self.radius = float(radius)
        self.device = _abi.require_device(device if device is not None else "cuda")
        w = ball_cell_weights(radius, grid.spacing, quad)
        self.weights = w
        self.m = (w.shape[0] - 1) // 2
        self.ball_volume = 4.0 / 3.0 * math.pi * radius ** 3
        self.quadrature_volume = float(w.sum()) * grid.spacing ** 3
        if any(2 * self.m + 1 > n for n in grid.shape):
            raise ValueError("ball kernel larger than the grid")
        handle = C.c_void_p()
        wc = np.ascontiguousarray(w, dtype=np.float64)
        if method not in ("auto", "direct", "fft"):
            raise ValueError("method must be 'auto', 'direct' or 'fft'")
        with _abi.device_lock:
            _abi.check(_abi.lib.svimx_nr_plan_create(wc.ctypes.data_as(C.c_void_p), self.m, float(grid.spacing),
                                                     self.radius, *map(int, grid.shape), C.byref(handle)),
                       "N_r plan")
            self.plan = handle
            self._ws = None
            if method != "auto":
                _abi.check(_abi.lib.svimx_nr_plan_set_method(handle, 1 if method == "direct" else 2), "N_r method")

    @property
    def method(self) -> str:
        """'direct' (summation at the needed nodes) or 'fft' (whole grid)."""
        return "fft" if _abi.lib.svimx_nr_plan_method(self.plan) == 2 else "direct"

    def __del__(self):
        plan = getattr(self, "plan", None)
        if plan is not None and _abi is not None and _abi.lib is not None:
            with _abi.device_lock:
                _abi.lib.svimx_nr_plan_destroy(plan)
            self.plan = None

    def _workspace(self):
        if self._ws is None:
            self._ws = _abi.workspace(self.grid.shape, self.device)
        return self._ws

    def exec_marked(self, occupancy: torch.Tensor, out: torch.Tensor, ws: torch.Tensor) -> None:
        """Evaluate N_r at every tile marked by the last fractions pass (no host sync)."""
        _abi.check(_abi.lib.svimx_nr_exec(_abi.stream(), self.plan, _abi.ptr(occupancy), _abi.ptr(out),
                                          _abi.ptr(ws), ws.numel()), "N_r")

    def exterior_fraction(self, fractions: torch.Tensor, nodes: torch.Tensor | None = None) -> torch.Tensor:
        """Dense N_r field for occupancy ``fractions``.

        With ``nodes`` (flat indices) only the tiles holding them are evaluated and
        all other entries are NaN; without, every node is evaluated.
        """
        f = _abi.occupancy_zeros(self.grid.shape, self.device)
        _abi.occupancy_view(f, self.grid.shape).copy_(fractions.to(self.device, torch.float32).view(self.grid.shape))
        out = torch.full(self.grid.shape, float("nan"), dtype=torch.float32, device=self.device)
        st = _abi.stream()
        _abi.check(_abi.lib.svimx_nr_reset(st, self.plan))
        if nodes is None:
            nodes = torch.arange(self.grid.size, dtype=torch.int32, device=self.device)
        nodes = nodes.to(self.device, torch.int32).contiguous()
        _abi.check(_abi.lib.svimx_nr_mark_nodes(st, self.plan, _abi.ptr(nodes), nodes.numel()))
        self.exec_marked(f, out, self._workspace())
        return out

    def exterior_fraction_at_points(self, fractions: torch.Tensor, points: torch.Tensor) -> torch.Tensor:
        """Dense N_r field, evaluated (at least) at every node that a tricubic
        interpolation at the index-space ``points`` (n, 3) reads; NaN elsewhere.
        The stencils are marked on the device, no node list is formed."""
        f = _abi.occupancy_zeros(self.grid.shape, self.device)
        _abi.occupancy_view(f, self.grid.shape).copy_(fractions.to(self.device, torch.float32).view(self.grid.shape))
        out = torch.full(self.grid.shape, float("nan"), dtype=torch.float32, device=self.device)
        p = points.to(self.device, torch.float32).reshape(-1, 3).contiguous()
        st = _abi.stream()
        _abi.check(_abi.lib.svimx_nr_reset(st, self.plan))
        _abi.check(_abi.lib.svimx_nr_mark_points(st, self.plan, _abi.ptr(p), p.shape[0]))
        self.exec_marked(f, out, self._workspace())
        return out

    @property
    def run_length(self) -> int:
        """Nodes per z-run of the direct summation (0: tile mode)."""
        return int(_abi.lib.svimx_nr_plan_run_length(self.plan))

    @run_length.setter
    def run_length(self, value: int) -> None:
        _abi.check(_abi.lib.svimx_nr_plan_set_run_length(self.plan, int(value)), "N_r run length")

    @property
    def lines(self) -> int:
        """y-lines per unit of the direct summation (1, 2 or 4; 0 in tile mode)."""
        return int(_abi.lib.svimx_nr_plan_lines(self.plan))

    @lines.setter
    def lines(self, value: int) -> None:
        _abi.check(_abi.lib.svimx_nr_plan_set_lines(self.plan, int(value)), "N_r lines per unit")

    @property
    def shape_name(self) -> str:
        """The summation shape as the tuner names it ("8", "8x2", "8/lds")."""
        return _shape_name(self.run_length, self.lines, self.staging == "lds")

    @property
    def staging(self) -> str:
        """How the direct summation reads the occupancy: "lds" (staged per wave
        through LDS) or "l1" (per-run windows)."""
        return "lds" if int(_abi.lib.svimx_nr_plan_staging(self.plan)) else "l1"

    @staging.setter
    def staging(self, value: str) -> None:
        if value not in ("lds", "l1"):
            raise ValueError(f"unknown N_r staging {value!r}")
        _abi.check(_abi.lib.svimx_nr_plan_set_staging(self.plan, 1 if value == "lds" else 0), "N_r staging")

    def tune(self, ls, cut_hi: float) -> dict:
        """Autotune the summation's shape (z-run length 4/8/16, y-lines per unit
        1/2/4) on the level set's current field; returns {shape name ("8",
        "8x2"): ms per launch} for the shapes this plan can take.  These
        shapes sum in one order, so the N_r values the velocity reads do not
        depend on the choice; the LDS-staged kernel (``staging = "lds"``) sums
        in another and is never picked.  Leaves the occupancy and N_r marks of
        ``ls``'s current field in place."""
        ms = (C.c_double * _TUNE_SHAPES)()
        field = torch.empty(ls.grid.size, dtype=torch.float32, device=self.device)
        _abi.check(_abi.lib.svimx_nr_plan_tune(_abi.stream(), self.plan, _abi.ptr(ls.flat), _abi.ptr(ls._band),
                                               ls.nb, *ls.dims, ls.h, float(cut_hi), _abi.ptr(ls.occupancy),
                                               _abi.ptr(field), _abi.ptr(ls.ws), ls.ws.numel(), ms), "N_r tuning")
        if self.run_length == 0:
            return {}
        return {_shape_name(*tune_shape(i)): ms[i] for i in range(_TUNE_SHAPES) if not math.isnan(ms[i])}

    def interior_volume(self, fractions: torch.Tensor, nodes: torch.Tensor | None = None) -> torch.Tensor:
        """Dense |B_r(x) cap Omega| (derived from N_r)."""
        return (1.0 - self.exterior_fraction(fractions, nodes).double()) * self.ball_volume
