"""Batch engine: many erosion jobs stepped together on one GPU.

Python face of ``svimx_engine_*`` (include/svimx.h).  A job is a prepared
:class:`~svierosion_amd.evolve.ErosionSimulation` (level set built, law and
N_r plan in place); each engine tick steps the running jobs its schedule selects
(all of them, or a priority order filled up to a band-node budget) with one
launch per kernel and RK stage for the whole selection, device-side dt and one
host synchronisation per tick.  Each job's arithmetic is that of the single-job
path (same device bodies, block decomposition and reduction order), so a job's
history and field are bit-identical alone, in any mix, and through
``ErosionSimulation.run``.  The step logic mirrors ErosionSimulation.run
(svierosion/evolve.py:182-250).

The simulation objects keep their device buffers; the engine owns only the
band-sized velocity / L0 arrays and the reinitialisation scratch.  After a
job stops, :meth:`BatchEngine.sync` writes the engine's state (current
buffer, band size, time, steps, history) back into the simulation, so meshing,
annotation and a later single-job ``run`` continue from it.
"""

from __future__ import annotations

import ctypes as C
import math

from . import _abi
from .evolve import ErosionSimulation, StepRecord
from .levelset import MARGIN

STOP_NAMES = {0: "running", 1: "target_volume", 2: "max_time", 3: "vanished", 4: "stalled", 5: "max_steps",
              6: "error"}
SCHEDULES = {"lockstep": 0, "fifo": 1, "longest": 2, "short": 3}


JobDesc, StepRec, JobStatus = _abi.JobDesc, _abi.StepRec, _abi.JobStatus


def job_desc(sim: ErosionSimulation) -> JobDesc:
    """The engine's view of a prepared simulation (pointers into its buffers)."""
    ls = sim.ls
    d = JobDesc()
    for i in range(3):
        d.phi[i] = ls.bufs[i].data_ptr()
    d.occupancy = ls.occupancy.data_ptr()
    d.band = ls._band.data_ptr()
    d.band_cap = ls._band.numel()
    d.ws = ls.ws.data_ptr()
    d.ws_bytes = ls.ws.numel()
    d.nx, d.ny, d.nz = ls.dims
    d.cur = ls.cur
    d.nb = ls.nb
    d.inside_outside = ls.inside_outside
    d.h = ls.h
    d.h3 = ls.h ** 3
    d.width = ls.width
    d.margin = MARGIN
    d.frac_valid = -1 if sim._frac_valid is None else int(sim._frac_valid)
    d.law = sim._law_abi
    if sim.conv is not None:
        d.nr_plan = sim.conv.plan
        d.field = sim.field.data_ptr()
        d.radius = sim.law.radius
    else:
        d.k12 = sim.k12.data_ptr()
    d.cfl = sim.cfl
    d.feedback_cfl = sim.feedback_cfl
    d.parabolic_cfl = sim.parabolic_cfl
    d.reinit_distance = sim.reinit_distance
    d.cutoff_lo, d.cutoff_hi = sim.cutoff
    d.reinit_iterations = sim.reinit_iterations
    d.time = sim.time
    d.initial_volume = sim.initial_volume
    d.displacement = sim._displacement
    d.steps = sim.steps
    return d


class BatchEngine:
    """Stepping of many :class:`ErosionSimulation` jobs in shared ticks (see module doc)."""

    def __init__(self, max_jobs: int = 64, graphs: bool = True):
        self._h = C.c_void_p()
        # the engine steps on a stream of its own (graph capture needs one that
        # is not the legacy default stream)
        _abi.check(_abi.lib.svimx_engine_create(None, int(max_jobs), C.byref(self._h)), "engine")
        self.sims: dict[int, ErosionSimulation] = {}
        self._synced: dict[int, int] = {}
        self._stop_buf, self._stop_n = (C.c_int * 256)(), C.c_int()
        self.use_graphs(graphs)

    def close(self) -> None:
        if self._h:
            _abi.lib.svimx_engine_destroy(self._h)
            self._h = C.c_void_p()
            self.sims.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def use_graphs(self, on: bool) -> None:
        _abi.check(_abi.lib.svimx_engine_use_graphs(self._h, int(bool(on))))

    def group_nodes(self, nodes: int) -> None:
        """Run each tick's jobs in launch groups of at most ``nodes`` band nodes (a larger
        job alone; 0: one group), each group through every stage before the next, so a
        large job's fields stay cached between its kernels.  Results do not depend on it."""
        _abi.check(_abi.lib.svimx_engine_group_nodes(self._h, int(nodes)), "engine groups")

    def schedule(self, policy: str = "lockstep", budget_nodes: int = 0, slack: float = 0.0,
                 short_work: float | None = None) -> None:
        """Which running jobs a tick steps: ``"lockstep"`` (all), or ``"fifo"`` (the order
        jobs were added) / ``"longest"`` (most estimated steps left first) filled while the
        selected band nodes stay within max(budget_nodes, (1 + slack) x the first job's);
        ``"short"``: the oldest job, joined by every job with at most ``short_work``
        estimated node-steps left, whatever its band (long jobs take their turn one
        after another)."""
        _abi.check(_abi.lib.svimx_engine_schedule(self._h, SCHEDULES[policy], int(budget_nodes), float(slack)),
                   "engine schedule")
        if short_work is not None:
            _abi.check(_abi.lib.svimx_engine_short_work(self._h, float(short_work)), "engine schedule")

    # ----------------------------------------------------------------- jobs
    def add(self, sim: ErosionSimulation) -> int:
        import torch
        if sim.newton_steps != 1:
            raise ValueError("the batch engine projects each node once (newton_steps = 1); "
                             "run this simulation with ErosionSimulation.run")
        torch.cuda.current_stream().synchronize()  # the job's buffers are ready
        desc = job_desc(sim)
        jid = C.c_int()
        _abi.check(_abi.lib.svimx_engine_add(self._h, C.byref(desc), C.byref(jid)), "engine add")
        self.sims[jid.value] = sim
        self._synced[jid.value] = 0
        return jid.value

    def run(self, job: int, target_fraction: float | None = 0.9, max_time: float | None = None,
            max_steps: int = 100000, tol: float = 1e-6) -> None:
        """Arm ``job`` with the stop conditions of ErosionSimulation.run."""
        _abi.check(_abi.lib.svimx_engine_run(self._h, job, -1.0 if target_fraction is None else float(target_fraction),
                                             -1.0 if max_time is None else float(max_time), int(max_steps),
                                             float(tol)), "engine run")

    def tick(self) -> int:
        """One step of each job the schedule selects; returns how many are still running."""
        n = C.c_int()
        with _abi.device_lock:  # no synchronous free elsewhere while a tick may capture
            rc = _abi.lib.svimx_engine_tick(self._h, C.byref(n))
        _abi.check(rc, "engine tick")
        return n.value

    def ticks(self, max_ticks: int) -> tuple[int, int]:
        """Up to ``max_ticks`` ticks, ending after one in which a job stopped; returns
        (jobs still running, ticks taken)."""
        n, k = C.c_int(), C.c_int()
        with _abi.device_lock:
            rc = _abi.lib.svimx_engine_ticks(self._h, int(max_ticks), C.byref(n), C.byref(k))
        _abi.check(rc, "engine ticks")
        return n.value, k.value

    def status(self, job: int) -> JobStatus:
        st = JobStatus()
        _abi.check(_abi.lib.svimx_engine_status(self._h, job, C.byref(st)))
        return st

    def stop_reason(self, job: int) -> str:
        return STOP_NAMES[self.status(job).stop_reason]

    def records(self, job: int, first: int = 0) -> list[StepRecord]:
        st = self.status(job)
        n = st.nrecords - first
        if n <= 0:
            return []
        buf = (StepRec * n)()
        _abi.check(_abi.lib.svimx_engine_records(self._h, job, first, n, buf))
        return [StepRecord(int(r.step), r.time, r.dt, r.volume, r.fraction, r.vmax, r.vmean, r.qmin, r.qmax,
                           int(r.capped), bool(r.reinit), r.wall) for r in buf]

    def sync(self, job: int) -> ErosionSimulation:
        """Write the engine's state of ``job`` back into its simulation."""
        sim = self.sims[job]
        st = self.status(job)
        ls = sim.ls
        ls.cur = int(st.cur)
        nb_changed = int(st.nb) != ls.nb
        ls.nb = int(st.nb)
        ls.inside_outside = int(st.inside_outside)
        sim.time = float(st.time)
        sim.steps = int(st.steps)
        sim._displacement = float(st.displacement)
        sim._frac_valid = None if st.frac_valid < 0 else int(st.frac_valid)
        sim.history.extend(self.records(job, self._synced[job]))
        self._synced[job] = int(st.nrecords)
        if nb_changed or sim.V.numel() < max(ls.nb, 1):
            sim._alloc_band_arrays()
        if st.stop_reason == 6:
            raise RuntimeError(f"engine job {job} failed: {st.message.decode(errors='replace')}")
        return sim

    def remove(self, job: int) -> None:
        _abi.check(_abi.lib.svimx_engine_remove(self._h, job))
        self.sims.pop(job, None)
        self._synced.pop(job, None)

    def stopped(self) -> list[int]:
        """Jobs that stopped since the last call (each reported once)."""
        buf, n = self._stop_buf, self._stop_n
        out = []
        while True:
            _abi.check(_abi.lib.svimx_engine_stopped(self._h, buf, len(buf), C.byref(n)))
            out.extend(buf[:n.value])
            if n.value < len(buf):
                return out

    def run_all(self) -> None:
        while self.tick() > 0:
            pass

    def timers_ms(self) -> dict:
        out = (C.c_double * 3)()
        _abi.check(_abi.lib.svimx_engine_timers(self._h, out))
        return {"tick": out[0], "secant": out[1], "reinit": out[2],
                "capture_fallbacks": int(_abi.lib.svimx_engine_capture_fallbacks(self._h))}


def run_result(engine: BatchEngine, job: int):
    """RunResult of the job's last run (after :meth:`BatchEngine.sync`)."""
    from .evolve import RunResult
    sim = engine.sims[job]
    st = engine.status(job)
    frac = float(st.volume) / sim.initial_volume if sim.initial_volume else math.nan
    return RunResult(STOP_NAMES[st.stop_reason], sim.steps, sim.time, frac, sim.history)
