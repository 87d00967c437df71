"""Erosion speed laws (API of svierosion/flows.py).

The law objects validate their parameters and compute the host-side constants
(``n_cap`` by bisection, svierosion/flows.py:71-80); the per-node speed and
its derivative run inside the HIP velocity kernel, which receives
``law.to_abi()``.  The torch ``speed`` / ``dspeed`` /
``speed_and_diffusivity`` methods evaluate the same formulas on tensors for
host-side annotation of output meshes.
"""

from __future__ import annotations

import torch

from . import _abi


class SpeedLaw:
    requires_svi = False
    requires_curvature = False
    name = "law"

    def describe(self) -> str:  # pragma: no cover
        return self.name

    def params(self) -> dict:  # pragma: no cover
        return {}


class SVIPowerLaw(SpeedLaw):
    """Inward speed V = N^k / (t - N)^s (+ eps), N = N_r (svierosion/flows.py:29-122)."""

    requires_svi = True

    def __init__(self, radius: float, k: int = 1, s: int = 0, t: float | None = None,
                 max_speed_ratio: float = 100.0, eps: float = 0.0):
        if radius <= 0:
            raise ValueError("radius must be positive")
        if int(k) != k or k < 0:
            raise ValueError("k must be a nonnegative integer")
        if s not in (0, 1):
            raise ValueError("s must be 0 or 1")
        if s == 1 and (t is None or not (0.5 < t < 1.0)):
            raise ValueError("s = 1 requires 1/2 < t < 1 (t must exceed N on flat parts, N = 1/2)")
        if eps < 0:
            raise ValueError("eps must be >= 0")
        self.radius = float(radius)
        self.k = int(k)
        self.s = int(s)
        self.t = None if t is None else float(t)
        self.eps = float(eps)
        self.max_speed_ratio = float(max_speed_ratio)
        self.n_cap = 1.0
        if self.s == 1:
            goal = self.max_speed_ratio * self._raw(0.5)
            lo, hi = 0.5, self.t
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                if self._raw(mid) < goal:
                    lo = mid
                else:
                    hi = mid
            self.n_cap = lo
        self.name = f"svi_k{self.k}_s{self.s}" + (f"_t{self.t:g}" if self.s else "")

    def _raw(self, n: float) -> float:
        v = n ** self.k
        if self.s:
            v /= (self.t - n)
        return v + self.eps

    def to_abi(self) -> _abi.Law:
        return _abi.Law(kind=0, k=self.k, s=self.s, reserved=0, t=self.t or 0.0, eps=self.eps,
                        n_cap=self.n_cap, alpha=1.0, delta=0.0)

    def speed(self, N: torch.Tensor) -> torch.Tensor:
        N = N.clamp(0.0, 1.0)
        if self.s:
            N = torch.clamp(N, max=self.n_cap)
            V = N ** self.k / (self.t - N)
        else:
            V = N ** self.k if self.k else torch.ones_like(N)
        return V + self.eps if self.eps else V

    def dspeed(self, N: torch.Tensor) -> torch.Tensor:
        N = N.clamp(1e-12, 1.0)
        if self.s:
            N = torch.clamp(N, max=self.n_cap)
            d = self.t - N
            return N ** self.k / d * (self.k / N + 1.0 / d)
        if self.k == 0:
            return torch.zeros_like(N)
        return self.k * N ** (self.k - 1)

    def describe(self) -> str:
        core = f"N^{self.k}/({self.t:g} - N)" if self.s else (f"N^{self.k}" if self.k else "1")
        if self.eps:
            core += f" + {self.eps:g}"
        return f"V = {core},  N = N_r with r = {self.radius:g}"

    def params(self) -> dict:
        return {"law": "svi", "radius": self.radius, "k": self.k, "s": self.s, "t": self.t,
                "eps": self.eps, "max_speed_ratio": self.max_speed_ratio, "n_cap": self.n_cap}


class _CurvatureLaw(SpeedLaw):
    requires_curvature = True
    _kind = 1

    def __init__(self, alpha: float = 1.0, eps: float = 0.0, delta: float = 0.0):
        if alpha <= 0:
            raise ValueError("alpha must be positive")
        if delta < 0:
            raise ValueError("delta must be >= 0")
        self.alpha = float(alpha)
        self.eps = float(eps)
        self.delta = float(delta)
        if self.alpha < 1 and self.delta == 0:
            raise ValueError("alpha < 1 needs a regularisation delta > 0 (d/dx x^alpha is infinite at 0, "
                             "which makes explicit time steps collapse)")

    def to_abi(self) -> _abi.Law:
        return _abi.Law(kind=self._kind, k=0, s=0, reserved=0, t=0.0, eps=self.eps, n_cap=1.0,
                        alpha=self.alpha, delta=self.delta)

    def _power(self, x):
        if self.alpha == 1.0:
            return x
        if self.alpha < 1:
            return x * (x + self.delta) ** (self.alpha - 1)
        return x ** self.alpha

    def _dpower(self, x):
        a = self.alpha
        if a == 1.0:
            return torch.ones_like(x)
        if a < 1:
            return (x + self.delta) ** (a - 2) * (a * x + self.delta)
        return a * x.clamp_min(0) ** (a - 1)


class MeanCurvatureLaw(_CurvatureLaw):
    """V = (H_+)^alpha + eps, H = (k1 + k2)/2 (svierosion/flows.py:161-183)."""

    name = "mean"
    _kind = 1

    def speed_and_diffusivity(self, k1, k2):
        H = 0.5 * (k1 + k2)
        Hp = H.clamp_min(0)
        V = self._power(Hp) + self.eps
        D = torch.where(H > 0, self._dpower(Hp), torch.zeros_like(H))
        return V, D

    def describe(self) -> str:
        return f"V = (H_+)^{self.alpha:g}" + (f" + {self.eps:g}" if self.eps else "")

    def params(self) -> dict:
        return {"law": "mean_curvature", "alpha": self.alpha, "eps": self.eps, "delta": self.delta}


class GaussCurvatureLaw(_CurvatureLaw):
    """V = ((k1)_+ (k2)_+)^alpha + eps (svierosion/flows.py:186-213)."""

    name = "gauss"
    _kind = 2

    def speed_and_diffusivity(self, k1, k2):
        a = k1.clamp_min(0)
        b = k2.clamp_min(0)
        Kc = a * b
        V = self._power(Kc) + self.eps
        dF = self._dpower(Kc)
        D = torch.where(Kc > 0, dF * (a + b), dF * 0)
        return V, D

    def describe(self) -> str:
        return f"V = (k1_+ k2_+)^{self.alpha:g}" + (f" + {self.eps:g}" if self.eps else "")

    def params(self) -> dict:
        return {"law": "gauss_curvature", "alpha": self.alpha, "eps": self.eps, "delta": self.delta}


def make_law(kind: str, **kw) -> SpeedLaw:
    kind = kind.lower()
    if kind == "svi":
        return SVIPowerLaw(**kw)
    if kind in ("mean", "mean_curvature", "h"):
        return MeanCurvatureLaw(**kw)
    if kind in ("gauss", "gauss_curvature", "k"):
        return GaussCurvatureLaw(**kw)
    raise ValueError(f"unknown law {kind!r}")


def flat_speed(law: SVIPowerLaw) -> float:
    """Speed of a flat surface (N = 1/2)."""
    return law._raw(0.5)


__all__ = ["SpeedLaw", "SVIPowerLaw", "MeanCurvatureLaw", "GaussCurvatureLaw", "make_law", "flat_speed"]
