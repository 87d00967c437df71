"""Batched and served erosion: many erosion jobs sharing one GPU.

The reference runs one ``erode`` at a time (svierosion/compare.py:36-85 loops
over runs sequentially).  Here a request loop does continuous batching over
the batch engine (svierosion_amd.engine):

* admission -- a request is admitted when it has arrived, an engine slot is
  free and its estimated device memory (:func:`job_device_bytes`) fits the
  HBM budget next to the jobs already resident;
* set-up and finishing -- worker threads, each on a stream of its own, build
  each admitted job (mesh SDF, level set, N_r plan: ``prepare_job``) and
  mesh and annotate each job that reached a target (``finish_target``),
  while the main thread keeps ticking the engine;
* stepping -- a tick advances the running jobs its schedule selects by one
  RK3 step each; jobs join and leave between ticks, so new requests start
  while others are running.  ``schedule="short"`` (the default of
  :func:`serve`) steps the oldest request in every tick and lets every
  request with little estimated work left (steps left x band nodes at most
  2e10 node-steps, or fewer than 4 steps in) run beside it: short requests
  no longer queue behind long ones, and long ones still run one after another
  in arrival order, as on a one-at-a-time queue, instead of sharing the GPU
  (which made early long requests finish last).  ``"fifo"`` takes jobs in
  arrival order and fills each tick with later ones while the selected band
  nodes stay within ``budget_nodes`` (or ``1 + slack`` times the first job's
  band).  ``"longest"`` puts the jobs with the most estimated steps left first;
  ``"lockstep"`` (the default of :func:`erode_batch`) steps every job in
  every tick, which finishes a batch soonest when one job needs far more
  steps than the rest: the others run beside it at full concurrency and the
  long job ends alone on the single-job path (12 jobs of the fragment mix:
  7.1 s lockstep, 7.8 s longest-first or FIFO).

Each job computes exactly what ``erode`` computes for it alone: the engine
reproduces the single-job path bit for bit, and set-up and meshing are the
same functions.  :func:`serve` returns per-request completions with arrival,
admission and finish times; :func:`serve_sequential` is the one-at-a-time
baseline over the same requests.
"""

from __future__ import annotations

import concurrent.futures as cf
import math
import sys
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .api import ErosionJob, ErosionParams, ErosionResult, erode, finish_target, job_result, prepare_job
from .meshio import TriMesh

ERROR_STOP = 6


@dataclass
class Request:
    mesh: TriMesh
    params: ErosionParams
    arrival: float = 0.0  # seconds after the start of serving
    tag: object = None


@dataclass
class Completion:
    index: int
    arrival: float
    admitted: float
    finished: float
    result: ErosionResult | None = None
    error: str | None = None
    device_bytes: int = 0
    steps: int = 0

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def job_device_bytes(mesh: TriMesh, params: ErosionParams) -> int:
    """Device memory of one job: the level set (3 dense field buffers, padded
    occupancy, band list, workspace), N_r field (SVI) or two curvature fields,
    the N_r plan's run marks and list, the engine's band arrays and the
    reinitialisation scratch, plus the SDF build's transient arrays."""
    from .api import make_grid
    g = make_grid(mesh, params)
    n = int(g.size)
    nx, ny, nz = (int(s) for s in g.shape)
    occ = 4 * int(_abi.lib.svimx_occupancy_elems(nx, ny, nz))
    ws = int(_abi.lib.svimx_workspace_bytes(nx, ny, nz))
    # band nodes: a shell of 2*band cells around the surface
    area = float(mesh.area()) if hasattr(mesh, "area") else 0.0
    nb = min(n, int(2.0 * params.band * area / g.spacing ** 2) + 4096) if area > 0 else n // 4
    svi = params.law == "svi"
    fields = 4 * n if svi else 8 * n
    plan = (n // 8) * 6 if svi else 0
    engine = 2 * 4 * nb + 2 * n + (2 * nb + 4096) * 40
    sdf_transient = 12 * n
    return 12 * n + occ + 4 * n + ws + fields + plan + engine + sdf_transient


def _quantile(xs, q):
    if not xs:
        return math.nan
    return float(np.quantile(np.asarray(xs, dtype=np.float64), q))


@dataclass
class ServeReport:
    completions: list
    makespan: float
    peak_device_bytes: int = 0
    ticks: int = 0
    engine_ms: dict = field(default_factory=dict)

    @property
    def done(self) -> list:
        return [c for c in self.completions if c.error is None]

    def erosions_per_hour(self) -> float:
        return 3600.0 * len(self.done) / self.makespan if self.makespan > 0 else 0.0

    def latency(self, q: float) -> float:
        return _quantile([c.latency for c in self.done], q)

    def summary(self) -> dict:
        return {"completed": len(self.done), "failed": len(self.completions) - len(self.done),
                "makespan_s": round(self.makespan, 3), "erosions_per_hour": round(self.erosions_per_hour(), 1),
                "latency_p50_s": round(self.latency(0.5), 3), "latency_p99_s": round(self.latency(0.99), 3),
                "peak_device_gb": round(self.peak_device_bytes / 1e9, 3), "ticks": self.ticks,
                "capture_fallbacks": int(self.engine_ms.get("capture_fallbacks", 0))}


class _Worker:
    """Host threads, each with its own torch stream, for set-up and meshing."""

    def __init__(self, threads: int = 2):
        self.pool = cf.ThreadPoolExecutor(max(1, int(threads)), thread_name_prefix="svimx-worker")
        self._local = threading.local()

    def _stream(self):
        import torch
        s = getattr(self._local, "stream", None)
        if s is None:
            s = self._local.stream = torch.cuda.Stream()
        return s

    def submit(self, fn, *args):
        import torch

        def call():
            with torch.cuda.stream(self._stream()):
                out = fn(*args)
                torch.cuda.current_stream().synchronize()
                return out
        return self.pool.submit(call)

    def close(self):
        self.pool.shutdown(wait=True)


# Tick budget of the `fifo` and `longest` schedules (`short`, the serving
# default, lets its short jobs join whatever their band), from a sweep of
# `fifo` over 12 Poisson requests of
# the served workload at 3.08/s (one-at-a-time FIFO: p50 1.19 s, p99 2.43 s,
# 8,750 erosions/hour): 2M band nodes / slack 0.5 gave p50 1.23 s, p99 2.01 s,
# 9,930 erosions/hour; slack 0 or a 1M budget lowered p50 to 1.20-1.24 s and
# raised p99 to 2.17-2.26 s; no budget (one job per tick) 1.35 / 2.62 s, the
# engine's per-step cost for a job alone being ~10% above the single-job path.
DEFAULT_BUDGET_NODES = 2_000_000
DEFAULT_SLACK = 0.5


def serve(requests: list, max_jobs: int = 64, hbm_budget: float | None = None, graphs: bool = True,
          keep_state: bool = False, verbose: bool = False, workers: int = 2, solo_tail: bool = True,
          schedule: str = "short", budget_nodes: int = DEFAULT_BUDGET_NODES, slack: float = DEFAULT_SLACK,
          ticks_per_visit: int = 8) -> ServeReport:
    """Continuous batching of ``requests`` (arrival times in seconds from now).  With
    ``solo_tail`` the last running job finishes on the single-job path once no
    request is left to admit.  ``schedule``, ``budget_nodes`` and ``slack`` select
    the jobs of each tick (:meth:`BatchEngine.schedule`); the loop looks at
    arrivals, set-up and meshing after every ``ticks_per_visit`` ticks or as soon
    as a job stops."""
    import torch
    from .engine import STOP_NAMES, BatchEngine
    _abi.require_device("cuda")
    if hbm_budget is None:
        free, _ = torch.cuda.mem_get_info()
        hbm_budget = 0.85 * free
    order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    pending = deque(order)
    eng = BatchEngine(max_jobs=max_jobs, graphs=graphs)
    eng.schedule(schedule, budget_nodes, slack)
    worker = _Worker(workers)
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0  # noqa: E731
    comps: dict[int, Completion] = {}
    prep: dict = {}      # future -> request index
    fin: dict = {}       # future -> (job id, request index)
    live: dict = {}      # job id -> (request index, ErosionJob, bytes, evolve start, steps at run start)
    evolve_s: dict = {}  # request index -> seconds in the engine
    used = 0
    peak = 0
    ticks = 0
    n_running = 0

    def complete(idx, result=None, error=None):
        c = comps[idx]
        c.finished = clock()
        c.result, c.error = result, error
        if verbose:
            print(f"[serve] request {idx} done at {c.finished:.2f}s "
                  f"({'ok' if error is None else error})", file=sys.stderr, flush=True)

    while pending or prep or fin or live:
        now = clock()
        # admission: arrived requests, FIFO, within the slot count and HBM budget
        while pending and requests[pending[0]].arrival <= now and len(live) + len(prep) < max_jobs:
            idx = pending[0]
            r = requests[idx]
            need = job_device_bytes(r.mesh, r.params)
            if used > 0 and used + need > hbm_budget:
                break
            pending.popleft()
            used += need
            peak = max(peak, used)
            comps[idx] = Completion(idx, r.arrival, now, math.nan, device_bytes=need)
            prep[worker.submit(prepare_job, r.mesh, r.params)] = idx
        # prepared jobs join the engine
        for fut in [f for f in prep if f.done()]:
            idx = prep.pop(fut)
            try:
                job = fut.result()
            except Exception as ex:  # parameter errors of erode (ValueError, RuntimeError)
                used -= comps[idx].device_bytes
                complete(idx, error=f"{type(ex).__name__}: {ex}")
                continue
            jid = eng.add(job.sim)
            eng.run(jid, job.targets[job.next_target], max_steps=job.params.max_steps)
            live[jid] = (idx, job, comps[idx].device_bytes, clock(), job.sim.steps)
            evolve_s[idx] = 0.0
            n_running += 1
        # the last running job, with nothing left to admit or mesh, leaves the
        # engine: a tick of one job costs more than a step of the single-job
        # path (every job-table kernel loads its tables first), and
        # ErosionSimulation.run takes the same steps bit for bit
        if solo_tail and n_running == 1 and not pending and not prep and not fin:
            jid = next(j for j in live if eng.status(j).running)
            idx, job, need, t_run, step0 = live.pop(jid)
            eng.sync(jid)
            eng.remove(jid)
            n_running = 0
            sim, err = job.sim, None
            try:
                res = sim.run(target_fraction=job.targets[job.next_target],
                              max_steps=job.params.max_steps - (sim.steps - step0))
                evolve_s[idx] += clock() - t_run
                while finish_target(job, res.stop_reason, res.fraction):
                    t_run = clock()
                    res = sim.run(target_fraction=job.targets[job.next_target], max_steps=job.params.max_steps)
                    evolve_s[idx] += clock() - t_run
            except Exception as ex:  # as in the meshing worker below
                err = f"{type(ex).__name__}: {ex}"
            comps[idx].steps = sim.steps
            if err is None:
                complete(idx, result=job_result(job, evolve_s[idx], keep_state))
            else:
                complete(idx, error=err)
            used -= need
            continue
        # up to ticks_per_visit ticks (fewer when a job stops), then look at
        # arrivals, set-up and meshing again
        if n_running > 0:
            n_running, k = eng.ticks(ticks_per_visit)
            if verbose and ticks // 5000 != (ticks + k) // 5000:
                print(f"[serve] {ticks + k} ticks at {clock():.2f}s, {n_running} jobs running", file=sys.stderr, flush=True)
            ticks += k
        # jobs that reached a target: mesh them on the worker
        for jid in eng.stopped():
            idx, job, need, t_run, _ = live[jid]
            evolve_s[idx] += clock() - t_run
            st = eng.status(jid)
            if st.stop_reason == ERROR_STOP:
                complete(idx, error=f"RuntimeError: {st.message.decode(errors='replace')}")
                eng.remove(jid)
                del live[jid]
                used -= need
                continue
            eng.sync(jid)
            fraction = float(st.volume) / job.sim.initial_volume
            fin[worker.submit(finish_target, job, STOP_NAMES[st.stop_reason], fraction)] = (jid, idx)
        for fut in [f for f in fin if f.done()]:
            jid, idx = fin.pop(fut)
            _, job, need, _, _ = live[jid]
            try:
                more = fut.result()
            except Exception as ex:
                more, err = False, f"{type(ex).__name__}: {ex}"
            else:
                err = None
            if more:
                eng.run(jid, job.targets[job.next_target], max_steps=job.params.max_steps)
                live[jid] = (idx, job, need, clock(), job.sim.steps)
                n_running += 1
                continue
            comps[idx].steps = job.sim.steps
            if err is None:
                complete(idx, result=job_result(job, evolve_s[idx], keep_state))
            else:
                complete(idx, error=err)
            eng.remove(jid)
            del live[jid]
            used -= need
        if n_running == 0:
            # nothing to step: wait for the worker or the next arrival
            waits = list(prep) + list(fin)
            nxt = requests[pending[0]].arrival - clock() if pending else None
            if waits:
                cf.wait(waits, timeout=None if nxt is None else max(nxt, 0.0), return_when=cf.FIRST_COMPLETED)
            elif nxt is not None and nxt > 0:
                time.sleep(nxt)
    makespan = clock()
    timers = eng.timers_ms()
    eng.close()
    worker.close()
    return ServeReport([comps[i] for i in sorted(comps)], makespan, int(peak), ticks, timers)


def serve_sequential(requests: list, keep_state: bool = False, verbose: bool = False,
                     wait_for_arrivals: bool = True) -> ServeReport:
    """The same requests served one at a time with ``erode`` (first come, first served).

    With ``wait_for_arrivals=False`` the requests run back to back and the FIFO
    schedule is replayed from the measured service times: request i starts at
    max(arrival_i, finish_{i-1}) and finishes a service time later -- the same
    single-server queue, without the idle waits for arrivals."""
    order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    t0 = time.perf_counter()
    comps = []
    clock = 0.0  # replayed FIFO clock
    for idx in order:
        r = requests[idx]
        if wait_for_arrivals:
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        s0 = time.perf_counter()
        c = Completion(idx, r.arrival, s0 - t0, math.nan)
        progress = None
        if verbose:  # a line every ~30 s inside long erosions
            last = [s0]

            def progress(sim, rec, idx=idx, last=last):
                now = time.perf_counter()
                if now - last[0] > 30.0:
                    last[0] = now
                    print(f"[sequential] request {idx}: step {rec.step}, fraction {rec.fraction:.5f}",
                          file=sys.stderr, flush=True)
        try:
            c.result = erode(r.mesh, r.params, keep_state=keep_state, callback=progress)
            c.steps = len(c.result.history)
        except Exception as ex:
            c.error = f"{type(ex).__name__}: {ex}"
        s1 = time.perf_counter()
        if wait_for_arrivals:
            c.finished = s1 - t0
        else:
            c.admitted = max(r.arrival, clock)
            c.finished = clock = c.admitted + (s1 - s0)
        comps.append(c)
        if verbose:
            print(f"[sequential] request {idx} done at {c.finished:.2f}s ({s1 - s0:.2f} s of service)",
                  file=sys.stderr, flush=True)
    comps.sort(key=lambda c: c.index)
    return ServeReport(comps, (time.perf_counter() - t0) if wait_for_arrivals else clock)


def erode_batch(meshes: list, params, max_jobs: int = 64, graphs: bool = True, keep_state: bool = False,
                hbm_budget: float | None = None, workers: int = 2, schedule: str = "lockstep",
                budget_nodes: int = DEFAULT_BUDGET_NODES, slack: float = DEFAULT_SLACK) -> list:
    """``erode`` of every mesh (one ErosionParams or one per mesh), batched on one GPU.

    Returns ErosionResult per mesh, or the exception a job raised."""
    plist = params if isinstance(params, (list, tuple)) else [params] * len(meshes)
    rep = serve([Request(m, p) for m, p in zip(meshes, plist)], max_jobs=max_jobs, hbm_budget=hbm_budget,
                graphs=graphs, keep_state=keep_state, workers=workers, schedule=schedule,
                budget_nodes=budget_nodes, slack=slack)
    return [c.result if c.error is None else RuntimeError(c.error) for c in rep.completions]
