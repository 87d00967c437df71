"""Time integration of the erosion flows on the device (API of svierosion/evolve.py).

Each RHS evaluation (one TVD-RK3 stage) is three libsvimx launches over the
band plus the N_r summation over the marked tiles:

    fractions (+ N_r tile marks)   (SVI laws)
    N_r direct summation           (SVI laws)   | curvature scatter (curvature laws)
    velocity fused with the RK stage (Godunov/WENO5, L = V |grad phi|, stage
    combination); on stage 0 it stores L0 with the dt statistics, and once the
    host has dt the update u0 + dt L0 is a short streaming pass

(``fused = False`` runs the velocity and the RK stage as two passes: the same
floats, bit-identical fields.)

Buffers: the step-start field u0 stays in one dense copy (A); stage outputs go
to B, C, B, so a secant retry restarts from A without a copy and replays
stage 0 from the saved L0 (svierosion/evolve.py:215-231).  dt is computed on
the device from stage 0's statistics (``svimx_stable_dt``, the host formula
bit for bit) and the later stages read it there, so the host reads the
results once per accepted step, after the volume of the new state; the
fractions of that state are reused by the next step's first stage.
"""

from __future__ import annotations

import contextlib
import math
import time as _time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi
from .flows import SpeedLaw
from .levelset import LevelSet
from .svi import BallConvolution


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


@dataclass
class RunResult:
    stop_reason: str
    steps: int
    time: float
    fraction: float
    history: list = field(default_factory=list)


class ErosionSimulation:
    """Evolve a :class:`LevelSet` under a speed law (svierosion/evolve.py:67-250)."""

    def __init__(self, levelset: LevelSet, law: SpeedLaw, cfl: float = 0.5, feedback_cfl: float = 1.0,
                 parabolic_cfl: float = 0.25, reinit_distance: float = 1.5, reinit_iterations: int = 20,
                 cutoff=(4.0, 5.5), newton_steps: int = 1, autotune: bool = True, nr_method: str = "auto"):
        self.ls = levelset
        self.law = law
        self.cfl = float(cfl)
        self.feedback_cfl = float(feedback_cfl)
        self.parabolic_cfl = float(parabolic_cfl)
        self.reinit_distance = float(reinit_distance)
        self.reinit_iterations = int(reinit_iterations)
        self.cutoff = tuple(float(c) for c in cutoff)
        if self.cutoff[1] > levelset.width:
            raise ValueError("cutoff must lie inside the band")
        if int(newton_steps) < 1:
            raise ValueError("newton_steps must be at least 1")
        self.newton_steps = int(newton_steps)
        self.cp = None  # newton_steps > 1: the band's closest points, 3 floats per entry
        self._law_abi = law.to_abi()
        dev = levelset.device
        n = levelset.grid.size
        self.conv = None
        self.field = None
        self.k12 = None  # curvature laws: dense (k1, k2) pairs, 2 floats per node
        if law.requires_svi:
            self.conv = BallConvolution(levelset.grid, law.radius, device=dev, method=nr_method)
            self._check_convolution_margin()
            self.field = torch.zeros(n, dtype=torch.float32, device=dev)
        else:
            self.k12 = torch.zeros(2 * n, dtype=torch.float32, device=dev)
        self._alloc_band_arrays()
        self._fv = None             # buffer index whose fractions/marks are current
        self._nr_valid = None       # buffer whose N_r (or curvature pairs) are current (None, or _frac_valid)
        self._res_host = None       # pinned landing area of the result slots
        self._res_event = None
        # run(): the velocity pass fused with the RK stage that consumes it
        # (svimx_velocity_rk); False runs svimx_velocity + svimx_rk_stage, the
        # same floats in two passes
        self.fused = True
        # run() with fused: dt computed on the device from stage 0's statistics
        # (svimx_stable_dt, bit-identical to stable_dt); False reads the
        # statistics on the host first (two host round trips per step)
        self.device_dt = True
        self.tuning = {}
        if self.conv is not None and autotune and levelset.nb > 0:
            # pick the N_r z-run length on this job's field (results do not depend on it)
            self.tuning = self.conv.tune(levelset, self.cutoff[1])
        self.initial_volume = self.ls.volume()
        self.time = 0.0
        self.steps = 0
        self.history: list[StepRecord] = []
        self._displacement = 0.0
        self._last_inc = 0.0        # front displacement dt max V of the last step
        self._reinit_error = None   # raised by a queued reinitialisation, re-raised in order
        self.timers = {"stats_sync": 0.0, "volume_sync": 0.0}
        # name -> [(start, end) timing-event pairs]; None disables event recording
        self.kernel_events = None
        # stage names to record (None: every stage); each span costs two event records
        self.kernel_event_names = None
        self.kernel_event_every = 1  # record every n-th span of a name
        self._span_calls = {}

    @property
    def _frac_valid(self):
        return self._fv

    @_frac_valid.setter
    def _frac_valid(self, buf) -> None:
        # new fractions/marks (or none) leave no N_r current
        self._fv = buf
        self._nr_valid = None

    @contextlib.contextmanager
    def _span(self, name: str):
        """Bracket the launches of one stage with HIP events on the current stream
        (every kernel_event_every-th span of each name)."""
        if self.kernel_events is None or (self.kernel_event_names is not None and name not in self.kernel_event_names):
            yield
            return
        k = self._span_calls.get(name, 0)
        self._span_calls[name] = k + 1
        if k % max(int(self.kernel_event_every), 1):
            yield
            return
        a = _abi.TimingEvent()
        b = _abi.TimingEvent()
        a.record()
        yield
        b.record()
        self.kernel_events.setdefault(name, []).append((a, b))

    def kernel_times_ms(self) -> dict:
        """Per-stage list of event durations (synchronises)."""
        torch.cuda.synchronize()
        return {k: [a.elapsed_time(b) for a, b in v] for k, v in (self.kernel_events or {}).items()}

    # ------------------------------------------------------------- checks
    def _check_convolution_margin(self) -> None:
        """Circular and direct sums agree if the solid stays m+1 cells from the boundary."""
        m = self.conv.m + 1
        inside = self.ls.phi < 0
        for ax in range(3):
            other = tuple(a for a in range(3) if a != ax)
            hit = torch.nonzero(inside.any(dim=other[1]).any(dim=other[0])).squeeze(1)
            if hit.numel() == 0:
                continue
            if int(hit.min()) < m or int(hit.max()) > self.ls.dims[ax] - 1 - m:
                raise ValueError("grid padding is smaller than the SVI ball radius; use a larger padding")

    def _alloc_band_arrays(self) -> None:
        """Band-indexed V and L0 (and the closest points) for at least nb entries;
        kept while large enough and grown with 25% headroom, so the band's drift
        at reinitialisations rarely reallocates them."""
        need = max(self.ls.nb, 1)
        if getattr(self, "V", None) is not None and self.V.numel() >= need and \
                (self.newton_steps == 1 or (self.cp is not None and self.cp.numel() >= 3 * need)):
            return
        cap = need + need // 4
        dev = self.ls.device
        self.V = self.L0 = self.cp = None
        self.V = torch.empty(cap, dtype=torch.float32, device=dev)
        self.L0 = torch.empty(cap, dtype=torch.float32, device=dev)
        if self.newton_steps > 1:
            self.cp = torch.empty(3 * cap, dtype=torch.float32, device=dev)

    # ---------------------------------------------------------- stages
    def _plan(self):
        return self.conv.plan if self.conv is not None else None

    def _inputs(self, buf: int) -> None:
        """What the velocity of the field in bufs[buf] reads: fractions and N_r
        (SVI laws) or the dense curvature pairs (curvature laws)."""
        ls = self.ls
        st = _abi.stream()
        saved = ls.cur
        ls.cur = buf
        try:
            # the occupancy feeds N_r only; curvature laws need the fractions
            # just for the volume at the end of the step
            if self.conv is not None and self._frac_valid != buf:
                with self._span("fractions"):
                    ls.fractions(self._plan(), self.cutoff[1])
                self._frac_valid = buf
            if self.newton_steps > 1:
                # the refined closest points of every band node (svierosion/levelset.py:
                # 139-145); their tricubic stencils join the N_r marks of the fractions pass
                _abi.check(_abi.lib.svimx_closest_points_newton(st, _abi.ptr(ls.flat), _abi.ptr(ls._band), ls.nb,
                                                                *ls.dims, ls.h, self.newton_steps,
                                                                _abi.ptr(self.cp)), "closest points")
                if self.conv is not None:
                    _abi.check(_abi.lib.svimx_nr_mark_points(st, self.conv.plan, _abi.ptr(self.cp), ls.nb),
                               "N_r marks")
            if self.conv is not None:
                if self._nr_valid != buf:
                    self._nr_inputs()
            elif self._nr_valid != buf:
                self._curvature_inputs()
                self._nr_valid = None  # the pairs no longer belong to a speculated buffer
        finally:
            ls.cur = saved

    def _nr_inputs(self) -> None:
        """N_r at the nodes the next velocity pass reads: selection and summation
        over the marks and occupancy of the last fractions pass."""
        ls = self.ls
        st = _abi.stream()
        with self._span("nr_select"):
            _abi.check(_abi.lib.svimx_nr_select(st, self.conv.plan), "N_r select")
        with self._span("nr_sum"):
            _abi.check(_abi.lib.svimx_nr_sum(st, self.conv.plan, _abi.ptr(ls.occupancy), _abi.ptr(self.field),
                                             _abi.ptr(ls.ws), ls.ws.numel()), "N_r")

    def _curvature_inputs(self) -> None:
        """The dense (k1, k2) pairs of the current field at the band nodes."""
        ls = self.ls
        with self._span("curvatures"):
            _abi.check(_abi.lib.svimx_curvatures(_abi.stream(), _abi.ptr(ls.flat), _abi.ptr(ls._band), ls.nb,
                                                 *ls.dims, ls.h, _abi.ptr(self.k12)))

    def _read_results(self, speculate: int | None = None, then=None):
        """Host copy of the result slots.  With ``speculate`` = the buffer whose
        fractions were just computed, that buffer's N_r (what the next step's
        stage 0 reads) is launched between the copy and the wait, so the device
        keeps working while the host reads the results.  ``then``: work queued
        between the copy and the wait instead; returns (results, its value)."""
        ls = self.ls
        if self._res_host is None:
            self._res_host = torch.empty(_abi.R_COUNT, dtype=torch.float64, pin_memory=True)
            self._res_event = torch.cuda.Event()
        self._res_host.copy_(_abi.results_view(ls.ws), non_blocking=True)
        self._res_event.record()
        out = None
        if then is not None:
            out = then()
        elif speculate is not None:
            if self.conv is not None:
                self._nr_inputs()
            else:
                saved = ls.cur
                ls.cur = speculate
                try:
                    self._curvature_inputs()
                finally:
                    ls.cur = saved
            self._nr_valid = speculate
        t0 = _time.perf_counter()
        self._res_event.synchronize()
        self.timers["volume_sync"] += _time.perf_counter() - t0
        r = self._res_host.numpy().copy()
        return r if then is None else (r, out)

    def _prepare(self, buf: int, stats: bool) -> None:
        """Fractions, N_r / curvatures and velocity (into self.V) for the field in bufs[buf]."""
        self._inputs(buf)
        with self._span("velocity"):
            self._velocity_call(_abi.stream(), self.ls.bufs[buf], stats)

    def _velocity_rk(self, stage: int, src: int, u0: int, dst: int, dt: float) -> None:
        """Fused velocity + RK stage on bufs[src] (inputs already prepared): stage 0
        stores L0 and the statistics, stages 1 and 2 write bufs[dst]."""
        ls = self.ls
        args = (_abi.stream(), stage, _abi.ptr(ls.bufs[src]), _abi.ptr(ls.bufs[u0]), _abi.ptr(ls.bufs[dst]),
                _abi.ptr(ls._band), ls.nb, *ls.dims, ls.h, self._law_abi, _abi.ptr(self.field), _abi.ptr(self.k12),
                self.cutoff[0], self.cutoff[1], float(dt), _abi.ptr(self.L0))
        with self._span("velocity_rk"):
            if self.newton_steps > 1:
                rc = _abi.lib.svimx_velocity_rk_points(*args, _abi.ptr(self.cp), _abi.ptr(ls.ws), ls.ws.numel())
            else:
                rc = _abi.lib.svimx_velocity_rk(*args, _abi.ptr(ls.ws), ls.ws.numel())
            _abi.check(rc, "velocity + rk stage")
        if stage > 0 and self._frac_valid == dst:
            self._frac_valid = None

    def _stage0(self, a: int) -> None:
        """Velocity with the dt statistics on bufs[a] (and L0 when fused)."""
        if self.fused:
            self._inputs(a)
            self._velocity_rk(0, a, a, a, 0.0)
        else:
            self._prepare(a, True)

    def _stage0_update(self, a: int, b: int, dt: float) -> None:
        """bufs[b] = u0 + dt L0 (the fused stage 0 left L0; the replay is the same update)."""
        self._stage(3 if self.fused else 0, a, a, b, dt)

    def _velocity_call(self, st, phi, stats: bool) -> None:
        ls = self.ls
        _abi.check(_abi.lib.svimx_velocity(st, _abi.ptr(phi), _abi.ptr(ls._band), ls.nb, *ls.dims, ls.h,
                                               self._law_abi, _abi.ptr(self.field), _abi.ptr(self.k12),
                                               self.cutoff[0], self.cutoff[1], int(stats),
                                               _abi.ptr(self.V), _abi.ptr(ls.ws), ls.ws.numel()), "velocity")

    def _stage(self, stage: int, src: int, u0: int, dst: int, dt: float) -> None:
        ls = self.ls
        with self._span("rk_stage"):
            self._stage_call(stage, src, u0, dst, dt)
        if self._frac_valid == dst:
            self._frac_valid = None

    def _stage_call(self, stage: int, src: int, u0: int, dst: int, dt: float) -> None:
        ls = self.ls
        args = (_abi.stream(), stage, _abi.ptr(ls.bufs[src]), _abi.ptr(ls.bufs[u0]), _abi.ptr(ls.bufs[dst]),
                _abi.ptr(ls._band), ls.nb, *ls.dims, ls.h)
        if dt == _abi.DT_DEVICE:
            # the step svimx_stable_dt left in the workspace's dt slot
            rc = _abi.lib.svimx_rk_stage_dt(*args, _abi.ptr_at(ls.ws, 8 * _abi.R_DT), _abi.ptr(self.V),
                                            _abi.ptr(self.L0))
        else:
            rc = _abi.lib.svimx_rk_stage(*args, float(dt), _abi.ptr(self.V), _abi.ptr(self.L0))
        _abi.check(rc, "rk stage")

    def _device_dt(self, max_dt: float | None) -> None:
        """stable_dt (and the max_time clamp) on the device from the statistics of
        the stage-0 pass, into the workspace's dt slot; bit-identical to the host
        formula."""
        ls = self.ls
        _abi.check(_abi.lib.svimx_stable_dt(_abi.stream(), self._law_abi, ls.h, self.cfl, self.feedback_cfl,
                                            self.parabolic_cfl, float(getattr(self.law, "radius", 0.0)),
                                            int(max_dt is not None), float(max_dt or 0.0), _abi.ptr(ls.ws),
                                            ls.ws.numel()), "stable dt")

    def _rk3_tail(self, a: int, b: int, c: int, dt: float) -> None:
        """Stages 1 and 2 after stage 0 wrote bufs[b]; the result lands in bufs[b]."""
        if self.fused:
            self._inputs(b)
            self._velocity_rk(1, b, a, c, dt)
            self._inputs(c)
            self._velocity_rk(2, c, a, b, dt)
            return
        self._prepare(b, False)
        self._stage(1, b, a, c, dt)
        self._prepare(c, False)
        self._stage(2, c, a, b, dt)

    def _volume_of(self, buf: int, speculate: bool = False) -> tuple[float, np.ndarray]:
        """Volume of bufs[buf] (its fractions stay current).  ``speculate``: the
        next step will most likely start from this buffer, so its N_r is launched
        before the host waits for the volume."""
        ls = self.ls
        saved = ls.cur
        ls.cur = buf
        try:
            with self._span("fractions"):
                ls.fractions(self._plan(), self.cutoff[1])
            self._frac_valid = buf
            spec = buf if speculate and self.newton_steps == 1 else None
            r = self._read_results(spec)
            return ls.volume_from_results(r), r
        finally:
            ls.cur = saved

    def _reinit_likely(self) -> bool:
        """Predict that the step in flight ends in a reinitialisation: its front
        displacement taken as 1.05 times the last step's reaches the threshold."""
        h = self.ls.h
        return self._displacement + 1.05 * self._last_inc >= self.reinit_distance * h

    def _zero_curvatures_on_band(self) -> None:
        # the dense curvature fields must read 0 off the band after a
        # reinitialisation, like the reference's freshly zeroed array
        # (svierosion/evolve.py:138); the next curvature pass rewrites the band
        if self.k12 is not None:
            self.k12.view(-1, 2).index_fill_(0, self.ls.band.long(), 0.0)

    def _volume_and_gated_reinit(self, buf: int, target: float | None, tol: float):
        """Volume of bufs[buf] with the reinitialisation of that field queued
        behind it (svimx_reinit_gated): it runs iff the step triggers it, decided
        on the device with the host's rule and operation order, so the host does
        not wait for the results before starting it.  Returns (volume, results,
        whether it ran)."""
        ls = self.ls
        io0 = ls.inside_outside
        gate = _abi.ReinitGate(displacement=self._displacement, threshold=self.reinit_distance * ls.h,
                               inside_outside=float(io0), h3=ls.h ** 3,
                               target=0.0 if target is None else target,
                               target_stop=0.0 if target is None else target * (1 + tol),
                               has_target=0 if target is None else 1)
        saved = ls.cur
        ls.cur = buf
        try:
            with self._span("fractions"):
                ls.fractions(self._plan(), self.cutoff[1])
            self._frac_valid = buf

            def reinit_behind():
                self._zero_curvatures_on_band()
                with self._span("reinit"):
                    try:
                        return ls.reinitialize(self.reinit_iterations, gate=gate)
                    except Exception as e:  # raised where the host's rule runs it
                        self._reinit_error = e
                        return True

            r, ran = self._read_results(then=reinit_behind)
            # the step's volume, over the band it was computed on
            return (io0 + float(r[_abi.R_VOLUME_SUM])) * ls.h ** 3, r, bool(ran)
        finally:
            ls.cur = saved

    def _gate_agrees(self, ran: bool, decided: bool) -> None:
        if ran and not decided:
            if self._reinit_error is not None:
                e, self._reinit_error = self._reinit_error, None
                raise e
            raise RuntimeError("the device-side reinitialisation trigger ran where the host's rule does not")

    def _no_reinit_after_step(self) -> bool:
        """True when the step in flight cannot trigger a reinitialisation: it moves
        the front by dt max V <= cfl h (dt <= cfl h / max V), so the displacement
        stays below the reinitialisation distance."""
        h = self.ls.h
        return self._displacement + self.cfl * h * (1.0 + 1e-9) < self.reinit_distance * h

    def _stats(self) -> tuple[float, dict]:
        t0 = _time.perf_counter()
        r = self.ls.results()
        self.timers["stats_sync"] += _time.perf_counter() - t0
        return self._stats_of(r)

    @staticmethod
    def _stats_of_law(law, r) -> tuple[float, dict]:
        info = {}
        # statistics over |phi| < h, or over every evaluated node when that set is
        # empty (svierosion/evolve.py:131, :147)
        near = r[_abi.R_NNEAR] > 0
        if law.requires_svi:
            if r[_abi.R_NALL] > 0 or near:
                info["qmin"] = float(r[_abi.R_QMIN if near else _abi.R_QMIN_ALL])
                info["qmax"] = float(r[_abi.R_QMAX if near else _abi.R_QMAX_ALL])
                info["dvmax"] = float(r[_abi.R_DVMAX if near else _abi.R_DVMAX_ALL])
                info["capped"] = int(r[_abi.R_CAPPED if near else _abi.R_CAPPED_ALL])
        else:
            info["qmin"] = float(r[_abi.R_QMIN if near else _abi.R_QMIN_ALL])
            info["qmax"] = float(r[_abi.R_QMAX if near else _abi.R_QMAX_ALL])
            info["dmax"] = float(r[_abi.R_DMAX if near else _abi.R_DMAX_ALL])
        if near:
            info["vmean"] = float(r[_abi.R_VSUM] / r[_abi.R_NNEAR])
        elif r[_abi.R_NALL] > 0:
            info["vmean"] = float(r[_abi.R_VSUM_ALL] / r[_abi.R_NALL])
        return float(r[_abi.R_VMAX]), info

    def _stats_of(self, r) -> tuple[float, dict]:
        """max V and the record's statistics from the result slots of stage 0."""
        return self._stats_of_law(self.law, r)

    def stable_dt(self, vmax: float, info: dict) -> float:
        """min(cfl h/max V, feedback_cfl (4r/3)/max dV/dN, parabolic_cfl h^2/max D) (svierosion/evolve.py:162-170)."""
        h = self.ls.h
        dt = self.cfl * h / vmax if vmax > 0 else math.inf
        if self.law.requires_svi and info.get("dvmax", 0) > 0:
            dt = min(dt, self.feedback_cfl * (4.0 * self.law.radius / 3.0) / info["dvmax"])
        if self.law.requires_curvature and info.get("dmax", 0) > 0:
            dt = min(dt, self.parabolic_cfl * h * h / info["dmax"])
        return dt

    def svi_field(self, nodes: torch.Tensor | None = None) -> torch.Tensor:
        """Dense N_r of the current shape (NaN outside the evaluated tiles when ``nodes`` is given)."""
        return self.conv.exterior_fraction(self.ls.dense_fractions(), nodes)

    def svi_field_at(self, points: torch.Tensor) -> torch.Tensor:
        """Dense N_r of the current shape, valid wherever a tricubic interpolation
        at the index-space ``points`` reads (NaN where nothing was evaluated)."""
        return self.conv.exterior_fraction_at_points(self.ls.dense_fractions(), points)

    # ---------------------------------------------------------------- run
    def run(self, target_fraction: float | None = 0.9, max_steps: int = 100000, max_time: float | None = None,
            tol: float = 1e-6, callback=None) -> RunResult:
        """Evolve until the volume equals ``target_fraction`` of the initial volume (or
        ``max_time`` / ``max_steps``); re-entrant with decreasing targets."""
        if self.newton_steps > 1 and not self.fused:
            raise ValueError("newton_steps > 1 runs on the fused velocity pass (fused = True)")
        ls = self.ls
        V0 = self.initial_volume
        target = None if target_fraction is None else target_fraction * V0
        ls.reserve_reinit()
        vol, res = self._volume_of(ls.cur, speculate=True)
        reason = "max_steps"
        n_start = self.steps
        while self.steps - n_start < max_steps:
            if target is not None and vol <= target * (1 + tol):
                reason = "target_volume"
                break
            if max_time is not None and self.time >= max_time * (1 - 1e-12):
                reason = "max_time"
                break
            if not (res[_abi.R_HAS_NEG] > 0 and res[_abi.R_HAS_POS] > 0):
                reason = "vanished"
                break
            t_wall = _time.time()
            a = ls.cur
            b, c = (a + 1) % 3, (a + 2) % 3
            pre_reinit = False
            self._stage0(a)
            if self.fused and self.device_dt:
                # dt on the device: the whole step runs without a host round trip
                # and the host reads the statistics, dt and the volume once
                self._device_dt(None if max_time is None else max_time - self.time)
                self._stage0_update(a, b, _abi.DT_DEVICE)
                self._rk3_tail(a, b, c, _abi.DT_DEVICE)
                if self._reinit_likely():
                    # queue the reinitialisation behind the step; the device runs
                    # it only if the step triggers it (the host's rule, below)
                    vnew, res, pre_reinit = self._volume_and_gated_reinit(b, target, tol)
                else:
                    vnew, res = self._volume_of(b, speculate=self._no_reinit_after_step())
                vmax, info = self._stats_of(res)
                dt = float(res[_abi.R_DT])
                if not math.isfinite(dt) or dt <= 0:
                    reason = "stalled"  # the step's stages applied 0; u0 stays current
                    self._gate_agrees(pre_reinit, False)
                    break
            else:
                vmax, info = self._stats()
                dt = self.stable_dt(vmax, info)
                if max_time is not None:
                    dt = min(dt, max_time - self.time)
                if not math.isfinite(dt) or dt <= 0:
                    reason = "stalled"
                    break
                self._stage0_update(a, b, dt)
                self._rk3_tail(a, b, c, dt)
                vnew, res = self._volume_of(b, speculate=self._no_reinit_after_step())
            if target is not None and vnew < target:
                self._gate_agrees(pre_reinit, False)
                lo_t, lo_v, hi_t, hi_v = 0.0, vol, dt, vnew
                for _ in range(8):
                    if hi_v >= lo_v:
                        break
                    dts = lo_t + (target - lo_v) * (hi_t - lo_t) / (hi_v - lo_v)
                    self._stage(3, a, a, b, dts)
                    self._rk3_tail(a, b, c, dts)
                    v, res = self._volume_of(b)
                    if abs(v - target) <= tol * V0:
                        break
                    if v > target:
                        lo_t, lo_v = dts, v
                    else:
                        hi_t, hi_v = dts, v
                dt, vnew = dts, v
            ls.cur = b
            self.time += dt
            self.steps += 1
            self._displacement += dt * vmax
            self._last_inc = dt * vmax
            did_reinit = False
            reached = target is not None and vnew <= target * (1 + tol)
            if self._displacement >= self.reinit_distance * ls.h and not reached:
                self._gate_agrees(pre_reinit, True)
                if self._reinit_error is not None:
                    e, self._reinit_error = self._reinit_error, None
                    raise e
                if not pre_reinit:
                    self._zero_curvatures_on_band()
                    with self._span("reinit"):
                        ls.reinitialize(self.reinit_iterations)
                self._alloc_band_arrays()
                self._frac_valid = None
                self._displacement = 0.0
                did_reinit = True
                vnew, res = self._volume_of(ls.cur, speculate=True)
            else:
                self._gate_agrees(pre_reinit, False)
            rec = StepRecord(self.steps, self.time, dt, vnew, vnew / V0, vmax, info.get("vmean", vmax),
                             info.get("qmin", float("nan")), info.get("qmax", float("nan")),
                             info.get("capped", 0), did_reinit, _time.time() - t_wall)
            self.history.append(rec)
            if callback is not None:
                callback(self, rec)
            vol = vnew
        return RunResult(reason, self.steps, self.time, vol / V0, self.history)
