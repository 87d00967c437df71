"""Erosion speed laws: the public law classes of svierosion/flows.py.

A law object here is a validated ``svimx_law_t`` record plus the public
attributes callers read (``radius``, ``k``, ``s``, ``t``, ``n_cap``, ``alpha``,
``delta``, ``eps``, ``requires_svi`` ...).  The library builds the record
(``svimx_law_init``: parameter checks and, for the geometric SVI form, the cap
n_cap), and the per-node formulas exist once, in the device code the velocity
pass runs.  ``speed``, ``dspeed`` and ``speed_and_diffusivity`` evaluate those
same device functions at given values (``svimx_op_law``), so what a caller
reads is what the evolution uses; they need the GPU like every other compute
entry point.
"""

from __future__ import annotations

import ctypes as C

import torch

from . import _abi

_SVI, _MEAN, _GAUSS = 0, 1, 2


def _record(kind: int, radius=0.0, k=0, s=0, t=None, eps=0.0, max_speed_ratio=100.0, alpha=1.0,
            delta=0.0) -> _abi.Law:
    rec = _abi.Law()
    tt = float("nan") if t is None else float(t)
    _abi.check(_abi.lib.svimx_law_init(int(kind), float(radius), float(k), float(s), tt, float(eps),
                                       float(max_speed_ratio), float(alpha), float(delta), C.byref(rec)))
    return rec


def _fmt(x: float) -> str:
    return f"{x:g}"


class SpeedLaw:
    """Base of the laws: the device record and evaluation at given values."""

    requires_svi = False
    requires_curvature = False
    name = "law"
    _rec: _abi.Law
    _param_keys: tuple = ()
    _param_tag = "law"

    def to_abi(self) -> _abi.Law:
        """A copy of the law's svimx_law_t record."""
        return _abi.Law.from_buffer_copy(self._rec)

    def _evaluate(self, q1: torch.Tensor, q2: torch.Tensor | None, want_d: bool):
        """Run svimx_op_law on the GPU; results in q1's dtype and device."""
        src = torch.as_tensor(q1)
        dev = _abi.require_device("cuda")
        a = src.to(device=dev, dtype=torch.float32).contiguous().reshape(-1)
        b = None if q2 is None else torch.as_tensor(q2).to(device=dev, dtype=torch.float32).contiguous().reshape(-1)
        V = torch.empty_like(a)
        D = torch.empty_like(a) if want_d else None
        _abi.check(_abi.lib.svimx_op_law(_abi.stream(), C.byref(self._rec), _abi.ptr(a), _abi.ptr(b), a.numel(),
                                         _abi.ptr(V), _abi.ptr(D)), "speed law")
        back = lambda x: x.reshape(src.shape).to(device=src.device, dtype=src.dtype if src.is_floating_point()
                                                  else torch.float32)
        return back(V), (back(D) if want_d else None)

    def params(self) -> dict:
        """The law's parameters as stored in result records (the reference's keys)."""
        return {"law": self._param_tag, **{key: getattr(self, key) for key in self._param_keys}}

    def describe(self) -> str:  # pragma: no cover - overridden
        return self.name


class SVIPowerLaw(SpeedLaw):
    """Inward speed V = N^k / (t - N)^s + eps of the exterior ball fraction N = N_r
    (svierosion/flows.py:29-122).  For s = 1, N is capped at n_cap, where the speed
    reaches ``max_speed_ratio`` times the flat-surface speed V(1/2)."""

    requires_svi = True
    _param_tag = "svi"
    _param_keys = ("radius", "k", "s", "t", "eps", "max_speed_ratio", "n_cap")

    def __init__(self, radius: float, k: int = 1, s: int = 0, t: float | None = None,
                 max_speed_ratio: float = 100.0, eps: float = 0.0):
        self._rec = _record(_SVI, radius=radius, k=k, s=s, t=t, eps=eps, max_speed_ratio=max_speed_ratio)
        self.radius = float(radius)
        self.k, self.s = int(self._rec.k), int(self._rec.s)
        self.t = None if t is None else float(t)
        self.eps = float(eps)
        self.max_speed_ratio = float(max_speed_ratio)
        self.n_cap = float(self._rec.n_cap)
        self.name = f"svi_k{self.k}_s{self.s}" + (f"_t{self.t:g}" if self.s else "")

    def _raw(self, n: float) -> float:
        """Uncapped speed at a scalar N in FP64 (host)."""
        out = C.c_double()
        _abi.check(_abi.lib.svimx_law_raw(C.byref(self._rec), float(n), C.byref(out)))
        return out.value

    def speed(self, N: torch.Tensor) -> torch.Tensor:
        return self._evaluate(N, None, False)[0]

    def dspeed(self, N: torch.Tensor) -> torch.Tensor:
        return self._evaluate(N, None, True)[1]

    def describe(self) -> str:
        num = "1" if self.k == 0 else ("N" if self.k == 1 else f"N^{self.k}")
        body = f"{num}/({_fmt(self.t)} - N)" if self.s else num
        tail = f" + {_fmt(self.eps)}" if self.eps else ""
        cap = f", N capped at {self.n_cap:.6f}" if self.s else ""
        return f"SVI power law V = {body}{tail} of N = N_r, ball radius {_fmt(self.radius)}{cap}"


class _Curvature(SpeedLaw):
    """Curvature laws V = F(q) + eps with F(x) = x^alpha, regularised for alpha < 1
    as x (x + delta)^(alpha - 1) (svierosion/flows.py:124-158)."""

    requires_curvature = True
    _kind = _MEAN
    _param_keys = ("alpha", "eps", "delta")
    _symbol = "q"

    def __init__(self, alpha: float = 1.0, eps: float = 0.0, delta: float = 0.0):
        self._rec = _record(self._kind, eps=eps, alpha=alpha, delta=delta)
        self.alpha, self.eps, self.delta = float(alpha), float(eps), float(delta)

    def speed_and_diffusivity(self, k1: torch.Tensor, k2: torch.Tensor):
        """(V, D) at principal curvatures (k1, k2), D = dV/dk1 + dV/dk2."""
        return self._evaluate(k1, k2, True)

    def describe(self) -> str:
        power = self._symbol if self.alpha == 1.0 else f"({self._symbol})^{_fmt(self.alpha)}"
        reg = f", regularised with delta = {_fmt(self.delta)}" if self.alpha < 1 else ""
        tail = f" + {_fmt(self.eps)}" if self.eps else ""
        return f"{self.name} curvature law V = {power}{tail}{reg}"


class MeanCurvatureLaw(_Curvature):
    """V = F(H_+) + eps with the mean curvature H = (k1 + k2)/2 (> 0 on a sphere)."""

    name = "mean"
    _kind = _MEAN
    _param_tag = "mean_curvature"
    _symbol = "H_+"


class GaussCurvatureLaw(_Curvature):
    """V = F((k1)_+ (k2)_+) + eps: the Gauss curvature at convex points, zero at
    saddles and concave points (alpha = 1/4 is the affine-invariant flow)."""

    name = "gauss"
    _kind = _GAUSS
    _param_tag = "gauss_curvature"
    _symbol = "k1_+ k2_+"


_ALIASES = {"svi": SVIPowerLaw, "mean": MeanCurvatureLaw, "mean_curvature": MeanCurvatureLaw,
            "h": MeanCurvatureLaw, "gauss": GaussCurvatureLaw, "gauss_curvature": GaussCurvatureLaw,
            "k": GaussCurvatureLaw}


def make_law(kind: str, **kw) -> SpeedLaw:
    """Law by name: svi, mean (mean_curvature, h) or gauss (gauss_curvature, k)."""
    cls = _ALIASES.get(str(kind).lower())
    if cls is None:
        raise ValueError(f"no speed law is called {kind!r}; use one of {sorted(_ALIASES)}")
    return cls(**kw)


def flat_speed(law: SVIPowerLaw) -> float:
    """Speed of a flat surface (N = 1/2) under an SVI law, in FP64."""
    return law._raw(0.5)


__all__ = ["SpeedLaw", "SVIPowerLaw", "MeanCurvatureLaw", "GaussCurvatureLaw", "make_law", "flat_speed"]
