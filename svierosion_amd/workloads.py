"""Seeded synthetic workloads of many erosion jobs.

* :func:`fragment_mix` -- the batched mix: fragments (icosphere level 4-7
  with spherical-harmonic radial displacement, radius 50), resolutions
  log-uniform in a range, laws 50% SVI (k in {0, 1, 2}, s in {0, 1},
  t = 0.93, r = 2.0), 25% mean curvature alpha = 1, 25% Gauss curvature
  alpha = 0.25 (delta automatic, 1/L^2), target volume fraction 0.9.
* :func:`poisson_requests` -- the served workload: the same fragments with
  resolutions log-uniform in [64, 400], targets uniform in [0.85, 0.97], and
  seeded Poisson arrivals at a fixed rate.

Job i's shape, law and resolution depend only on (seed, i), so a prefix of a
workload is the same jobs as the prefix of a longer one.
"""

from __future__ import annotations

import math

import numpy as np

from .api import ErosionParams
from .shapes import sh_fragment

SVI_RADIUS = 2.0


def _job(seed: int, i: int, res_lo: float, res_hi: float, target):
    rng = np.random.default_rng([seed, i])
    level = int(rng.integers(4, 8))
    res = int(round(math.exp(rng.uniform(math.log(res_lo), math.log(res_hi)))))
    u = rng.uniform()
    tf = float(target(rng)) if callable(target) else float(target)
    if u < 0.5:
        k = int(rng.integers(0, 3))
        s = int(rng.integers(0, 2))
        p = ErosionParams(law="svi", radius=SVI_RADIUS, k=k, s=s, t=0.93, resolution=res, target_fraction=tf)
    elif u < 0.75:
        p = ErosionParams(law="mean", alpha=1.0, resolution=res, target_fraction=tf)
    else:
        p = ErosionParams(law="gauss", alpha=0.25, resolution=res, target_fraction=tf)
    mesh = sh_fragment(level=level, radius=50.0, seed=seed * 100003 + i)
    return mesh, p


def fragment_mix(n: int = 64, seed: int = 4, res_range=(96, 320), target: float = 0.9):
    """[(mesh, ErosionParams)] of the batched fragment mix."""
    return [_job(seed, i, res_range[0], res_range[1], target) for i in range(n)]


def poisson_requests(n: int = 512, rate: float = 1.0, seed: int = 5, res_range=(64, 400)):
    """[Request] of the served workload; ``rate`` arrivals per second."""
    from .batch import Request
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        mesh, p = _job(seed, i, res_range[0], res_range[1], lambda r: r.uniform(0.85, 0.97))
        out.append(Request(mesh, p, float(arrivals[i]), tag=i))
    return out
