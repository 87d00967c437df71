"""Oracle (test infrastructure only): N_r, speed laws, level set and time stepping.

NumPy restatement of svierosion/svi.py, svierosion/flows.py,
svierosion/levelset.py and svierosion/evolve.py.  The dense field phi is a
flat float32 (or float64) array; the band is a sorted int64 index array.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import stencils as S

MARGIN = 4


# ------------------------------------------------------------------ ball weights / convolution
def ball_cell_weights(radius: float, spacing: float, quad: int = 32) -> np.ndarray:
    """Exact |B_r cap cell| / h^3 on the (2m+1)^3 offsets, m = floor(R+1/2)+1.

    Cells wholly inside/outside are classified exactly; cut cells integrate the
    z-extent analytically and (x, y) with quad x quad Gauss-Legendre points.
    svierosion/svi.py:24-55.
    """
    R = radius / spacing
    m = int(math.floor(R + 0.5)) + 1
    ax = np.arange(-m, m + 1, dtype=np.float64)
    I, J, K = np.meshgrid(ax, ax, ax, indexing="ij")
    cen = np.stack([I, J, K], -1)
    nearest = np.sqrt((np.maximum(np.abs(cen) - 0.5, 0) ** 2).sum(-1))
    farthest = np.sqrt(((np.abs(cen) + 0.5) ** 2).sum(-1))
    w = np.zeros(I.shape)
    w[farthest <= R] = 1.0
    cut = (nearest < R) & (farthest > R)
    g, gw = np.polynomial.legendre.leggauss(quad)
    g, gw = 0.5 * g, 0.5 * gw
    cc = cen[cut]
    X = cc[:, 0, None, None] + g[None, :, None]
    Y = cc[:, 1, None, None] + g[None, None, :]
    zr = np.sqrt(np.maximum(R * R - X * X - Y * Y, 0.0))
    zlo = cc[:, 2, None, None] - 0.5
    zhi = cc[:, 2, None, None] + 0.5
    L = np.maximum(np.minimum(zhi, zr) - np.maximum(zlo, -zr), 0.0)
    w[cut] = (L * gw[None, :, None] * gw[None, None, :]).sum((1, 2))
    return w


def ball_kernel_hat(shape, w):
    """rfftn of the weights placed at wrapped offsets.  svierosion/svi.py:66-82."""
    m = (w.shape[0] - 1) // 2
    if any(2 * m + 1 > n for n in shape):
        raise ValueError("ball kernel larger than the grid")
    kern = np.zeros(shape, dtype=np.float32)
    r = np.arange(-m, m + 1)
    kern[np.ix_(r % shape[0], r % shape[1], r % shape[2])] = w.astype(np.float32)
    return np.fft.rfftn(kern)


def exterior_fraction_fft(fractions, khat, h, radius):
    """N_r = 1 - h^3 (w * f) / |B_r| as a circular FFT convolution.  svierosion/svi.py:84-93."""
    shape = fractions.shape
    g = np.fft.irfftn(np.fft.rfftn(fractions) * khat, s=shape, axes=(0, 1, 2))
    bv = 4.0 / 3.0 * math.pi * radius ** 3
    return (1.0 - g * h ** 3 / bv).astype(fractions.dtype)


def exterior_fraction_direct(fractions, w, h, radius, points):
    """Brute-force direct sum of the same convolution at selected nodes (FP64).

    Independent check of the FFT form: N_r(x) = 1 - h^3 sum_d w(d) f(x+d) / |B_r|.
    ``points`` is an (n, 3) integer array; f is zero outside the grid.
    """
    m = (w.shape[0] - 1) // 2
    f = np.pad(fractions.astype(np.float64), m)
    out = np.empty(len(points))
    for q, (i, j, k) in enumerate(points):
        out[q] = (w * f[i:i + 2 * m + 1, j:j + 2 * m + 1, k:k + 2 * m + 1]).sum()
    bv = 4.0 / 3.0 * math.pi * radius ** 3
    return 1.0 - out * h ** 3 / bv


def sphere_exterior_fraction(r, R):
    """Closed-form N_r on a sphere.  svierosion/svi.py:96-101."""
    if r > 2 * R:
        return 1.0 - (R / r) ** 3
    return 0.5 + 3.0 * r / (16.0 * R)


# ------------------------------------------------------------------ speed laws
class SVILaw:
    """V = N^k/(t-N)^s + eps with the n_cap speed cap.  svierosion/flows.py:29-122."""

    requires_svi, requires_curvature = True, False

    def __init__(self, radius, k=1, s=0, t=None, max_speed_ratio=100.0, eps=0.0):
        if radius <= 0:
            raise ValueError("radius must be positive")
        if int(k) != k or k < 0:
            raise ValueError("k must be a nonnegative integer")
        if s not in (0, 1):
            raise ValueError("s must be 0 or 1")
        if s == 1 and (t is None or not (0.5 < t < 1.0)):
            raise ValueError("s = 1 requires 1/2 < t < 1")
        if eps < 0:
            raise ValueError("eps must be >= 0")
        self.radius, self.k, self.s = float(radius), int(k), int(s)
        self.t = None if t is None else float(t)
        self.eps, self.max_speed_ratio = float(eps), float(max_speed_ratio)
        self.n_cap = 1.0
        if self.s == 1:
            goal = self.max_speed_ratio * self.raw(0.5)
            lo, hi = 0.5, self.t
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                if self.raw(mid) < goal:
                    lo = mid
                else:
                    hi = mid
            self.n_cap = lo

    def raw(self, n):
        v = n ** self.k
        if self.s:
            v /= (self.t - n)
        return v + self.eps

    def speed(self, N):
        N = np.clip(N, 0.0, 1.0)
        if self.s:
            N = np.minimum(N, N.dtype.type(self.n_cap))
            V = N ** self.k / (N.dtype.type(self.t) - N)
        else:
            V = N ** self.k if self.k else np.ones_like(N)
        return V + N.dtype.type(self.eps) if self.eps else V

    def dspeed(self, N):
        N = np.clip(N, 1e-12, 1.0)
        if self.s:
            N = np.minimum(N, N.dtype.type(self.n_cap))
            d = N.dtype.type(self.t) - N
            V = N ** self.k / d
            return V * (self.k / N + 1.0 / d)
        if self.k == 0:
            return np.zeros_like(N)
        return self.k * N ** (self.k - 1)


class CurvatureLaw:
    """Mean (H+)^alpha or Gauss (k1+ k2+)^alpha, regularised for alpha < 1.

    svierosion/flows.py:124-213.
    """

    requires_svi, requires_curvature = False, True

    def __init__(self, kind, alpha=1.0, eps=0.0, delta=0.0):
        if alpha <= 0:
            raise ValueError("alpha must be positive")
        if delta < 0:
            raise ValueError("delta must be >= 0")
        if alpha < 1 and delta == 0:
            raise ValueError("alpha < 1 needs a regularisation delta > 0")
        self.kind, self.alpha, self.eps, self.delta = kind, float(alpha), float(eps), float(delta)

    def _power(self, x):
        a = self.alpha
        if a == 1.0:
            return x
        if a < 1:
            return x * (x + x.dtype.type(self.delta)) ** x.dtype.type(a - 1)
        return x ** x.dtype.type(a)

    def _dpower(self, x):
        a = self.alpha
        if a == 1.0:
            return np.ones_like(x)
        if a < 1:
            y = x + x.dtype.type(self.delta)
            return y ** x.dtype.type(a - 2) * (x.dtype.type(a) * x + x.dtype.type(self.delta))
        return x.dtype.type(a) * np.maximum(x, 0) ** x.dtype.type(a - 1)

    def speed_and_diffusivity(self, k1, k2):
        if self.kind == "mean":
            H = 0.5 * (k1 + k2)
            Hp = np.maximum(H, 0)
            V = self._power(Hp) + k1.dtype.type(self.eps)
            D = np.where(H > 0, self._dpower(Hp), 0).astype(k1.dtype)
            return V, D
        a = np.maximum(k1, 0)
        b = np.maximum(k2, 0)
        Kc = a * b
        V = self._power(Kc) + k1.dtype.type(self.eps)
        dF = self._dpower(Kc)
        D = np.where(Kc > 0, dF * (a + b), dF * 0).astype(k1.dtype)
        return V, D


# ------------------------------------------------------------------ level set
class LevelSet:
    """Dense phi + sorted band of |phi| < width*h.  svierosion/levelset.py:51-183."""

    def __init__(self, phi, shape, h, width=6.0):
        self.shape = tuple(shape)
        self.h = float(h)
        self.width = float(width)
        self.flat = np.ascontiguousarray(phi, dtype=phi.dtype).reshape(-1).copy()
        self.strides = S.strides_of(self.shape)
        self.rebuild_band()

    def rebuild_band(self):
        dt = self.flat.dtype.type
        cap = self.width * self.h
        np.clip(self.flat, dt(-cap), dt(cap), out=self.flat)
        self.band = np.nonzero(np.abs(self.flat) < dt(cap * (1 - 1e-5)))[0]
        self.check_margin(self.band)

    def check_margin(self, idx):
        if idx.size == 0:
            return
        ijk = S.unravel(idx, self.shape)
        for a in range(3):
            if ijk[:, a].min() < MARGIN or ijk[:, a].max() > self.shape[a] - 1 - MARGIN:
                raise RuntimeError("narrow band reaches the grid boundary; enlarge the padding")

    def has_interface(self):
        b = self.flat[self.band]
        return bool((b < 0).any()) and bool((b > 0).any())

    def interface_nodes(self):
        return self.band[S.interface_mask(self.flat, self.band, self.strides)]

    def cell_fractions(self, idx=None):
        idx = self.band if idx is None else idx
        return S.cell_fractions(self.flat, idx, self.strides)

    def dense_fractions(self):
        f = (self.flat < 0).astype(self.flat.dtype)
        f[self.band] = self.cell_fractions()
        return f.reshape(self.shape)

    def volume(self):
        f = self.flat
        out_band_inside = int((f < 0).sum()) - int((f[self.band] < 0).sum())
        return (out_band_inside + float(self.cell_fractions().astype(np.float64).sum())) * self.h ** 3

    def closest_points(self, idx, newton_steps=1):
        return S.closest_points(self.flat, idx, self.shape, self.h, newton_steps)

    def reinitialize(self, iterations=12, cfl=0.45, dilation=3, geometry64=False):
        """Min (2010) subcell-fixed PDE reinitialisation.  svierosion/levelset.py:154-183.

        geometry64: form the crossings, subcell distances and local dtau in
        float64 from the field's values and round them to the field's dtype once
        (the HIP path's precision form); the reference forms them in the field's
        dtype."""
        mask = np.zeros(self.flat.size, dtype=bool)
        mask[self.band] = True
        mask = dilate(mask.reshape(self.shape), dilation)
        for a in range(3):
            sl = [slice(None)] * 3
            sl[a] = slice(0, MARGIN)
            mask[tuple(sl)] = False
            sl[a] = slice(self.shape[a] - MARGIN, self.shape[a])
            mask[tuple(sl)] = False
        idx = np.nonzero(mask.reshape(-1))[0]
        f = self.flat
        if geometry64 and f.dtype != np.float64:
            geo = S.reinit_geometry(f.astype(np.float64), idx, self.strides, self.h, cfl)
            geo = [g.astype(f.dtype) if g.dtype == np.float64 else g for g in geo]
        else:
            geo = S.reinit_geometry(f, idx, self.strides, self.h, cfl)
        dtau = geo[1]
        for _ in range(iterations):
            u0 = f[idx].copy()
            f[idx] = u0 + dtau * S.reinit_rhs(f, idx, self.strides, self.h, geo)
            u1 = f[idx].copy()
            f[idx] = 0.5 * (u0 + u1 + dtau * S.reinit_rhs(f, idx, self.strides, self.h, geo))
        self.rebuild_band()


def dilate(mask, steps=1):
    """(2*steps+1)^3 cube dilation by separable passes.  svierosion/levelset.py:25-34."""
    out = mask.copy()
    for a in range(3):
        for _ in range(steps):
            src = out.copy()
            hi = [slice(None)] * 3
            lo = [slice(None)] * 3
            hi[a] = slice(1, None)
            lo[a] = slice(0, -1)
            out[tuple(hi)] |= src[tuple(lo)]
            out[tuple(lo)] |= src[tuple(hi)]
    return out


# ------------------------------------------------------------------ simulation
@dataclass
class StepRecord:
    step: int
    time: float
    dt: float
    volume: float
    fraction: float
    vmax: float
    vmean: float
    qmin: float = float("nan")
    qmax: float = float("nan")
    capped: int = 0
    reinit: bool = False
    wall: float = 0.0


class Simulation:
    """TVD-RK3 evolution with CFL control, secant stop and reinit trigger.

    svierosion/evolve.py:67-250.
    """

    def __init__(self, ls: LevelSet, law, cfl=0.5, feedback_cfl=1.0, parabolic_cfl=0.25,
                 reinit_distance=1.5, reinit_iterations=20, cutoff=(4.0, 5.5), newton_steps=1):
        self.ls, self.law = ls, law
        self.newton_steps = int(newton_steps)
        self.cfl, self.feedback_cfl, self.parabolic_cfl = float(cfl), float(feedback_cfl), float(parabolic_cfl)
        self.reinit_distance, self.reinit_iterations = float(reinit_distance), int(reinit_iterations)
        self.cutoff = tuple(float(c) for c in cutoff)
        if self.cutoff[1] > ls.width:
            raise ValueError("cutoff must lie inside the band")
        self.weights = None
        if law.requires_svi:
            self.weights = ball_cell_weights(law.radius, ls.h)
            self.m = (self.weights.shape[0] - 1) // 2
            self.khat = ball_kernel_hat(ls.shape, self.weights)
            self._check_convolution_margin()
        self.initial_volume = ls.volume()
        self.time, self.steps, self.history, self._disp = 0.0, 0, [], 0.0

    def _check_convolution_margin(self):
        m = self.m + 1
        inside = np.nonzero(self.ls.flat < 0)[0]
        ijk = S.unravel(inside, self.ls.shape)
        for a in range(3):
            if ijk[:, a].min() < m or ijk[:, a].max() > self.ls.shape[a] - 1 - m:
                raise ValueError("grid padding is smaller than the SVI ball radius")

    def svi_field(self):
        return exterior_fraction_fft(self.ls.dense_fractions(), self.khat, self.ls.h, self.law.radius)

    def _cutoff(self, phi_b):
        dt = phi_b.dtype.type
        a, b = dt(self.cutoff[0] * self.ls.h), dt(self.cutoff[1] * self.ls.h)
        x = np.abs(phi_b)
        t = np.clip((b - x) / (b - a), 0, 1)
        return t * t * (3 - 2 * t)

    def velocity(self, stats=False):
        ls = self.ls
        idx = ls.band
        phi_b = ls.flat[idx]
        cp = ls.closest_points(idx, self.newton_steps)
        info = {}
        near = np.abs(phi_b) < phi_b.dtype.type(ls.h)
        if self.law.requires_svi:
            Nd = self.svi_field()
            N = S.tricubic(Nd.reshape(-1), ls.shape, cp)
            V = self.law.speed(N)
            if stats:
                Nn = N[near] if near.any() else N
                info["qmin"], info["qmax"] = float(Nn.min()), float(Nn.max())
                info["dvmax"] = float(self.law.dspeed(Nn).max())
                info["capped"] = int((Nn > self.law.n_cap).sum())
        else:
            k1, k2 = S.principal_curvatures(ls.flat, idx, ls.strides, ls.h)
            dense = np.zeros(ls.flat.size, dtype=ls.flat.dtype)
            dense[idx] = k1
            k1c = S.trilinear(dense, ls.shape, cp)
            dense[idx] = k2
            k2c = S.trilinear(dense, ls.shape, cp)
            V, D = self.law.speed_and_diffusivity(k1c, k2c)
            if stats:
                sel = near if near.any() else np.ones_like(near)
                Hn = (0.5 * (k1c + k2c))[sel]
                info["qmin"], info["qmax"] = float(Hn.min()), float(Hn.max())
                info["dmax"] = float(D[sel].max())
        if stats:
            info["vmean"] = float(V[near].mean()) if near.any() else float(V.mean())
        return V * self._cutoff(phi_b), info

    def rhs(self, stats=False):
        V, info = self.velocity(stats)
        G = S.godunov_inward(self.ls.flat, self.ls.band, self.ls.strides, 1.0 / self.ls.h)
        return V * G, V, info

    def stable_dt(self, V, info):
        h = self.ls.h
        vmax = float(V.max()) if V.size else 0.0
        dt = self.cfl * h / vmax if vmax > 0 else math.inf
        if self.law.requires_svi and info.get("dvmax", 0) > 0:
            dt = min(dt, self.feedback_cfl * (4.0 * self.law.radius / 3.0) / info["dvmax"])
        if self.law.requires_curvature and info.get("dmax", 0) > 0:
            dt = min(dt, self.parabolic_cfl * h * h / info["dmax"])
        return dt

    def _rk3(self, u0, L0, dt):
        f, idx = self.ls.flat, self.ls.band
        d = f.dtype.type(dt)
        f[idx] = u0 + d * L0
        L1, _, _ = self.rhs()
        f[idx] = 0.75 * u0 + 0.25 * (f[idx] + d * L1)
        L2, _, _ = self.rhs()
        f[idx] = u0 / 3.0 + (2.0 / 3.0) * (f[idx] + d * L2)

    def run(self, target_fraction=0.9, max_steps=100000, max_time=None, tol=1e-6, callback=None):
        ls = self.ls
        V0 = self.initial_volume
        target = None if target_fraction is None else target_fraction * V0
        vol = ls.volume()
        reason = "max_steps"
        n0 = self.steps
        while self.steps - n0 < max_steps:
            if target is not None and vol <= target * (1 + tol):
                reason = "target_volume"
                break
            if max_time is not None and self.time >= max_time * (1 - 1e-12):
                reason = "max_time"
                break
            if not ls.has_interface():
                reason = "vanished"
                break
            idx = ls.band
            u0 = ls.flat[idx].copy()
            L0, Vb, info = self.rhs(stats=True)
            dt = self.stable_dt(Vb, info)
            if max_time is not None:
                dt = min(dt, max_time - self.time)
            if not math.isfinite(dt) or dt <= 0:
                reason = "stalled"
                break
            self._rk3(u0, L0, dt)
            vnew = ls.volume()
            if target is not None and vnew < target:
                lo_t, lo_v, hi_t, hi_v = 0.0, vol, dt, vnew
                for _ in range(8):
                    if hi_v >= lo_v:
                        break
                    dts = lo_t + (target - lo_v) * (hi_t - lo_t) / (hi_v - lo_v)
                    ls.flat[idx] = u0
                    self._rk3(u0, L0, dts)
                    v = ls.volume()
                    if abs(v - target) <= tol * V0:
                        break
                    if v > target:
                        lo_t, lo_v = dts, v
                    else:
                        hi_t, hi_v = dts, v
                dt, vnew = dts, v
            vmax = float(Vb.max())
            self.time += dt
            self.steps += 1
            self._disp += dt * vmax
            did = False
            reached = target is not None and vnew <= target * (1 + tol)
            if self._disp >= self.reinit_distance * ls.h and not reached:
                ls.reinitialize(self.reinit_iterations)
                self._disp = 0.0
                did = True
                vnew = ls.volume()
            rec = StepRecord(self.steps, self.time, dt, vnew, vnew / V0, vmax, info.get("vmean", vmax),
                             info.get("qmin", float("nan")), info.get("qmax", float("nan")),
                             info.get("capped", 0), did)
            self.history.append(rec)
            if callback is not None:
                callback(self, rec)
            vol = vnew
        return reason, self.steps, self.time, vol / V0, self.history
