"""Benchmark: RK3 steps/s of the narrow-band SVI erosion on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload (default): the 1.3M-triangle synthetic scan -- icosphere level 8
(1,310,720 triangles) scaled to semi-axes 60/45/30 with 4-octave seeded value
noise (amplitudes 3, 1.5, 0.75, 0.375), SVI law k=2 s=1 t=0.95, ball radius 2.0,
resolution 400 (grid 432x336x243).  The mesh SDF and level set are built
before the timed region; the timed region is K accepted RK3 steps (three RHS
evaluations each, with reinitialisation when triggered), bracketed by device
synchronisation.  Per-stage HIP events recorded on the launch stream give the
live kernel durations used for the roofline object.  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP32_VALU_PEAK_TFLOPS = 157.3  # packed FMA; unpacked is half


def build_workload(args):
    import torch
    from svierosion_amd import shapes
    from svierosion_amd.api import ErosionParams, initial_max_nr, make_grid
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.levelset import LevelSet
    from svierosion_amd.sdf import mesh_to_sdf

    t0 = time.time()
    mesh = shapes.noisy_ellipsoid(level=args.level, axes=(60.0, 45.0, 30.0),
                                  amplitudes=(3.0, 1.5, 0.75, 0.375), base_frequency=2.0, seed=args.seed)
    params = ErosionParams(law="svi", radius=2.0, k=2, s=1, t=0.95, resolution=args.resolution,
                           target_fraction=0.9, snapshots=(0.97, 0.94))
    lo, hi = mesh.bounds
    law = params.make_law(length_scale=float((hi - lo).max()))
    grid = make_grid(mesh, params)
    t1 = time.time()
    phi = mesh_to_sdf(mesh, grid, band=params.band)
    ls = LevelSet(phi, grid, width=params.band)
    sim = ErosionSimulation(ls, law, cfl=params.cfl, feedback_cfl=params.feedback_cfl,
                            reinit_distance=params.reinit_distance, reinit_iterations=params.reinit_iterations)
    nmax = initial_max_nr(sim)
    torch.cuda.synchronize()
    t2 = time.time()
    if nmax >= law.t:
        raise SystemExit(f"workload invalid: max N_r {nmax:.4f} >= t {law.t}")
    info = {"triangles": int(len(mesh.faces)), "grid": list(grid.shape), "nodes": int(grid.size),
            "band_nodes": int(ls.nb), "mesh_s": round(t1 - t0, 2), "sdf_setup_s": round(t2 - t1, 2),
            "nmax0": round(nmax, 5), "m": int(sim.conv.m), "nr_shape": sim.conv.shape_name,
            "nr_tuning_ms": {str(k): round(v, 4) for k, v in sim.tuning.items()}}
    return mesh, grid, ls, sim, law, info


def nr_read_bytes(sim) -> float:
    """Unique occupancy bytes the N_r summation needs: the active tiles dilated by m."""
    import torch
    from svierosion_amd import _abi
    plan = sim.conv
    ls = sim.ls
    # re-derive the tiles of the current state on the side (outside the timed region)
    ls.fractions(plan.plan, sim.cutoff[1])
    sim._frac_valid = None
    st = _abi.stream()
    _abi.check(_abi.lib.svimx_nr_select(st, plan.plan))
    out = torch.full(ls.dims, float("nan"), device=ls.device)
    _abi.check(_abi.lib.svimx_nr_sum(st, plan.plan, _abi.ptr(ls.occupancy), _abi.ptr(out), _abi.ptr(ls.ws),
                                     ls.ws.numel()))
    written = ~torch.isnan(out)
    n_out = int(written.sum())
    m = plan.m
    dil = written.float()[None, None]
    k = 2 * m + 1
    for ks, pad in (((k, 1, 1), (m, 0, 0)), ((1, k, 1), (0, m, 0)), ((1, 1, k), (0, 0, m))):
        dil = torch.nn.functional.max_pool3d(dil, kernel_size=ks, stride=1, padding=pad)
    n_in = int((dil[0, 0] > 0).sum())
    tiles = int(ls.results()[_abi.R_TILES])
    return n_out, n_in, tiles


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ls, law, steps=6):
    """The C++/OpenMP float32 restatement of the step (oracle/cpu/svicpu.cpp) on all the
    host cores this process may use: `steps` full accepted steps (three RHS
    evaluations with N_r by direct summation at the needed nodes, FP64 volume,
    reinitialisation when the displacement triggers it) from the same field."""
    from oracle import cpu_port
    from oracle.flow import SVILaw
    phi = ls.flat.detach().cpu().numpy().reshape(ls.dims).copy()
    ref_law = SVILaw(law.radius, law.k, law.s, law.t)
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    t0 = time.perf_counter()
    _, rows, threads = cpu_port.run_steps(phi, ls.h, ref_law, steps)
    el = time.perf_counter() - t0
    reinits = sum(1 for r in rows if r[2])
    return {"value": round(steps / el, 5), "unit": "RK3 steps/s", "cores": int(threads), "kind": "port",
            "cpu": _cpu_model(), "affinity_cores": cores,
            "sample": (f"{steps} full steps ({reinits} reinitialisations) of the same "
                       f"{ls.dims[0]}x{ls.dims[1]}x{ls.dims[2]} field, C++/OpenMP float32 restatement "
                       f"(oracle/cpu/svicpu.cpp) on {threads} threads: {el:.1f} s")}


def write_records(path, leg: str, reqs, report) -> None:
    """Append one JSON line per job of a batched or served leg: arrival,
    admission and finish times, latency, steps, law and grid."""
    if not path:
        return
    with open(path, "a") as fh:
        for c in report.completions:
            r = reqs[c.index]
            res = c.result
            fh.write(json.dumps({"leg": leg, "index": c.index, "law": r.params.law, "k": r.params.k,
                                 "s": r.params.s, "resolution": r.params.resolution,
                                 "target": r.params.target_fraction, "triangles": int(len(r.mesh.faces)),
                                 "grid": list(res.grid.shape) if res is not None else None,
                                 "arrival": round(c.arrival, 4), "admitted": round(c.admitted, 4),
                                 "finished": round(c.finished, 4), "latency": round(c.latency, 4),
                                 "steps": c.steps, "error": c.error,
                                 "timings": {k: round(v, 4) for k, v in res.timings.items()} if res else None})
                     + "\n")


def _same_history(a, b) -> bool:
    return len(a) == len(b) and all(x.volume == y.volume and x.dt == y.dt and x.reinit == y.reinit
                                    for x, y in zip(a, b))


def batched_leg(n_jobs: int, workers: int, schedule: str, budget_nodes: int, slack: float, progress: bool = False,
                records: str | None = None):
    """The fragment mix (first n_jobs of the 64-fragment set) eroded batched and one at a time."""
    from svierosion_amd.batch import Request, serve, serve_sequential
    from svierosion_amd.workloads import fragment_mix
    t0 = time.time()
    jobs = fragment_mix(n_jobs)
    reqs = [Request(m, p) for m, p in jobs]
    gen_s = time.time() - t0
    rb = serve(reqs, workers=workers, schedule=schedule, budget_nodes=budget_nodes, slack=slack, verbose=progress)
    print(f"[bench] batched {n_jobs} jobs: {rb.summary()}", file=sys.stderr, flush=True)
    write_records(records, "batched", reqs, rb)
    rs = serve_sequential(reqs, verbose=progress)
    print(f"[bench] sequential {n_jobs} jobs: {rs.summary()}", file=sys.stderr, flush=True)
    write_records(records, "batched_sequential", reqs, rs)
    same = all(cb.error is None and cs.error is None and _same_history(cb.result.history, cs.result.history)
               for cb, cs in zip(rb.completions, rs.completions))
    laws = {}
    for _, p in jobs:
        laws[p.law] = laws.get(p.law, 0) + 1
    steps = [c.steps for c in rb.completions]
    return {"workload": (f"first {n_jobs} of the 64-fragment mix (icosphere level 4-7 + SH displacement, "
                         "res log-uniform 96-320, 50% SVI / 25% mean / 25% Gauss curvature, target 0.9)"),
            "jobs": n_jobs, "laws": laws, "steps_min_max": [min(steps), max(steps)],
            "schedule": {"policy": schedule, "budget_nodes": budget_nodes, "slack": slack},
            "erosions_per_hour": round(rb.erosions_per_hour(), 1),
            "sequential_erosions_per_hour": round(rs.erosions_per_hour(), 1),
            "speedup": round(rb.erosions_per_hour() / max(rs.erosions_per_hour(), 1e-9), 3),
            "makespan_s": round(rb.makespan, 3), "sequential_makespan_s": round(rs.makespan, 3),
            "identical_histories": bool(same), "peak_device_gb": round(rb.peak_device_bytes / 1e9, 3),
            "engine_ticks": rb.ticks, "mesh_generation_s": round(gen_s, 2)}, rs


def served_leg(n_req: int, rate: float, workers: int, schedule: str, budget_nodes: int, slack: float,
               progress: bool = False, records: str | None = None, replay_fifo: bool = False):
    """Poisson arrivals of the served workload (first n_req requests), continuous batching vs FIFO
    (live, or with replay_fifo the requests back to back and the FIFO queue replayed from their
    measured service times)."""
    from svierosion_amd.batch import serve, serve_sequential
    from svierosion_amd.workloads import poisson_requests
    reqs = poisson_requests(n_req, rate=rate)
    rv = serve(reqs, workers=workers, schedule=schedule, budget_nodes=budget_nodes, slack=slack, verbose=progress)
    print(f"[bench] served {n_req} requests at {rate:.4f}/s: {rv.summary()}", file=sys.stderr, flush=True)
    write_records(records, "served", reqs, rv)
    rq = serve_sequential(reqs, verbose=progress, wait_for_arrivals=not replay_fifo)
    print(f"[bench] sequential FIFO: {rq.summary()}", file=sys.stderr, flush=True)
    write_records(records, "served_sequential", reqs, rq)
    same = all(a.error is None and b.error is None and _same_history(a.result.history, b.result.history)
               for a, b in zip(rv.completions, rq.completions))
    return {"workload": (f"first {n_req} of 512 requests (fragments as in the mix, res log-uniform 64-400, "
                         f"target U[0.85, 0.97]), seeded Poisson arrivals at {rate:.4f}/s"),
            "requests": n_req, "rate_per_s": round(rate, 5),
            "schedule": {"policy": schedule, "budget_nodes": budget_nodes, "slack": slack},
            "erosions_per_hour": round(rv.erosions_per_hour(), 1),
            "latency_p50_s": round(rv.latency(0.5), 3), "latency_p99_s": round(rv.latency(0.99), 3),
            "sequential_erosions_per_hour": round(rq.erosions_per_hour(), 1),
            "sequential_latency_p50_s": round(rq.latency(0.5), 3),
            "sequential_latency_p99_s": round(rq.latency(0.99), 3),
            "speedup": round(rv.erosions_per_hour() / max(rq.erosions_per_hour(), 1e-9), 3),
            "identical_histories": bool(same), "peak_device_gb": round(rv.peak_device_bytes / 1e9, 3),
            "fifo": "replayed from back-to-back service times" if replay_fifo else "live"}


def served_part(args) -> int:
    """One part of a long served measurement, for request counts whose two legs do not
    fit one run: ``--serve-part served`` runs the continuous-batching leg over the first
    --serve-requests requests; ``--serve-part fifo:A:B`` erodes requests A..B-1 back to
    back (their FIFO service times).  Both append per-request records to --records;
    --summarize-records combines the parts (the FIFO queue replayed in arrival order
    from the service times)."""
    from svierosion_amd.batch import ServeReport, serve, serve_sequential
    from svierosion_amd.workloads import poisson_requests
    from svierosion_amd.batch import DEFAULT_BUDGET_NODES, DEFAULT_SLACK
    if not args.records or not args.serve_rate:
        raise SystemExit("--serve-part needs --records and --serve-rate")
    reqs = poisson_requests(args.serve_requests, rate=args.serve_rate)
    if args.serve_part == "served":
        budget = DEFAULT_BUDGET_NODES if args.budget_nodes is None else args.budget_nodes
        slack = DEFAULT_SLACK if args.slack is None else args.slack
        rv = serve(reqs, workers=args.workers, schedule=args.serve_schedule, budget_nodes=budget, slack=slack,
                   verbose=args.progress)
        write_records(args.records, "served", reqs, rv)
        print(json.dumps({"serve_part": "served", "rate_per_s": args.serve_rate, **rv.summary()}), flush=True)
        return 0
    _, a, b = args.serve_part.split(":")
    part = [reqs[i] for i in range(int(a), min(int(b), len(reqs)))]
    rq = serve_sequential(part, verbose=args.progress, wait_for_arrivals=False)
    for c in rq.completions:  # back to the workload's request indices
        c.index += int(a)
    write_records(args.records, "served_sequential", reqs, rq)
    print(json.dumps({"serve_part": args.serve_part, "requests": len(part),
                      "service_s": round(sum(c.finished - c.admitted for c in rq.completions), 3)}), flush=True)
    return 0


def summarize_records(paths) -> int:
    """Served vs FIFO from per-request records (any number of files): the served
    leg's latencies as measured; the FIFO leg replayed in arrival order from each
    request's measured service time (start = max(arrival, previous finish)); the
    histories compared by final volume and step count.  One JSON line."""
    rows = [json.loads(l) for p in paths for l in open(p) if l.strip()]
    sv = {r["index"]: r for r in rows if r["leg"] == "served"}
    sq = {r["index"]: r for r in rows if r["leg"] == "served_sequential"}
    common = sorted(set(sv) & set(sq))
    clock, fifo = 0.0, {}
    for i in sorted(common, key=lambda i: sq[i]["arrival"]):
        start = max(sq[i]["arrival"], clock)
        clock = start + (sq[i]["finished"] - sq[i]["admitted"])
        fifo[i] = clock - sq[i]["arrival"]
    lat_s = np.array([sv[i]["latency"] for i in common])
    lat_f = np.array([fifo[i] for i in common])
    mk_s = max(sv[i]["finished"] for i in common)
    laws = {}
    for i in common:
        laws[sv[i]["law"]] = laws.get(sv[i]["law"], 0) + 1
    print(json.dumps({"requests": len(common), "laws": laws,
                      "steps_total": int(sum(sv[i]["steps"] for i in common)),
                      "served": {"makespan_s": round(mk_s, 2), "erosions_per_hour": round(3600 * len(common) / mk_s, 1),
                                 "latency_p50_s": round(float(np.quantile(lat_s, 0.5)), 3),
                                 "latency_p99_s": round(float(np.quantile(lat_s, 0.99)), 3)},
                      "fifo": {"makespan_s": round(clock, 2), "erosions_per_hour": round(3600 * len(common) / clock, 1),
                               "latency_p50_s": round(float(np.quantile(lat_f, 0.5)), 3),
                               "latency_p99_s": round(float(np.quantile(lat_f, 0.99)), 3)},
                      "same_steps": all(sv[i]["steps"] == sq[i]["steps"] for i in common)}), flush=True)
    return 0


def full_erosion(args) -> int:
    """The public erode() end to end (SDF, evolution to the target volume with
    its snapshots, meshing) on one of the single-job workloads: one JSON line
    with the phase times and ms per accepted step (evolution seconds / steps,
    reinitialisation and secant retries included, like
    benchmarks/performance.py:57)."""
    import torch
    from svierosion_amd import shapes
    from svierosion_amd.api import ErosionParams, erode
    if args.full_erosion == "cube64":
        mesh = shapes.cube(1.0)
        params = ErosionParams(law="svi", radius=0.1, k=1, s=1, t=0.93, resolution=64, target_fraction=0.9)
        name = "cube of side 1 (12 triangles), SVI k=1 s=1 t=0.93, r=0.1, res 64, to 0.9"
    elif args.full_erosion == "fragment256":
        mesh = shapes.sh_fragment(level=6, radius=50.0, seed=args.seed)
        params = ErosionParams(law="svi", radius=2.0, k=1, s=1, t=0.93, resolution=256, target_fraction=0.9)
        name = "icosphere(6) + 16 SH terms (81,920 triangles), SVI k=1 s=1 t=0.93, r=2.0, res 256, to 0.9"
    else:
        mesh = shapes.noisy_ellipsoid(level=args.level, axes=(60.0, 45.0, 30.0),
                                      amplitudes=(3.0, 1.5, 0.75, 0.375), base_frequency=2.0, seed=args.seed)
        params = ErosionParams(law="svi", radius=2.0, k=2, s=1, t=0.95, resolution=args.resolution,
                               target_fraction=0.9, snapshots=(0.97, 0.94))
        name = (f"{len(mesh.faces)}-triangle noisy ellipsoid scan, SVI k=2 s=1 t=0.95 r=2.0, res "
                f"{args.resolution}, to 0.9 with snapshots 0.97/0.94")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = erode(mesh, params)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    t = res.timings
    steps = len(res.history)
    line = {"full_erosion": args.full_erosion, "workload": name, "triangles": int(len(mesh.faces)),
            "grid": list(res.grid.shape), "stop_reason": res.stop_reason, "steps": steps,
            "reinits": sum(1 for r in res.history if r.reinit), "final_fraction": round(res.fraction, 7),
            "snapshots": [round(f, 4) for f, _, _ in res.snapshots],
            "sdf_s": round(t.get("sdf", 0.0), 3), "evolve_s": round(t.get("evolve", 0.0), 3),
            "mesh_s": round(t.get("mesh", 0.0), 3), "total_s": round(total, 3),
            "ms_per_step": round(1000 * t.get("evolve", 0.0) / max(steps, 1), 3)}
    print(json.dumps(line), flush=True)
    return 0


def nr_shape_sweep(sim) -> int:
    """Time each N_r summation shape the plan can take on the current field:
    one JSON line per shape (for per-shape profiler counters)."""
    import torch
    from svierosion_amd import _abi
    from svierosion_amd.svi import _TUNE_SHAPES, tune_shape
    conv, ls = sim.conv, sim.ls
    st = _abi.stream()
    out = torch.empty(ls.grid.size, dtype=torch.float32, device=ls.device)
    for i in range(_TUNE_SHAPES):
        rl, lines, lds = tune_shape(i)
        try:
            conv.staging, conv.lines = "l1", 1
            conv.run_length = rl
            conv.lines = lines
            conv.staging = "lds" if lds else "l1"
        except ValueError:
            continue
        ls.fractions(conv.plan, sim.cutoff[1])
        _abi.check(_abi.lib.svimx_nr_select(st, conv.plan))
        args = (st, conv.plan, _abi.ptr(ls.occupancy), _abi.ptr(out), _abi.ptr(ls.ws), ls.ws.numel())
        for _ in range(2):
            _abi.check(_abi.lib.svimx_nr_sum(*args))
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(5):
            _abi.check(_abi.lib.svimx_nr_sum(*args))
        b.record()
        torch.cuda.synchronize()
        units = int(ls.results()[_abi.R_TILES])
        print(json.dumps({"shape": conv.shape_name, "ms": round(a.elapsed_time(b) / 5, 4), "units": units,
                          "outputs": units * rl * max(lines, 1)}), flush=True)
    return 0


def nr_fft_timing(ls, sim, reps: int = 10) -> int:
    """The reference's N_r method (FFT over the whole grid, svierosion/svi.py:84-93) on
    the workload's occupancy: ms per evaluation (HIP events on the library's stream
    over `reps` evaluations) and the largest difference from the direct summation at
    the nodes the step reads.  One JSON line."""
    import torch
    from svierosion_amd import _abi
    from svierosion_amd.svi import BallConvolution
    g = ls.grid
    fft = BallConvolution(g, sim.conv.radius, device=ls.device, method="fft")
    st = _abi.stream()
    ls.fractions(sim.conv.plan, sim.cutoff[1])  # occupancy of the current field, and the step's N_r marks
    occ = ls.occupancy
    out = torch.empty(g.size, dtype=torch.float32, device=ls.device)
    ws = fft._workspace()
    nodes = torch.arange(g.size, dtype=torch.int32, device=ls.device)
    _abi.check(_abi.lib.svimx_nr_reset(st, fft.plan))
    _abi.check(_abi.lib.svimx_nr_mark_nodes(st, fft.plan, _abi.ptr(nodes), nodes.numel()))
    for _ in range(2):
        fft.exec_marked(occ, out, ws)
    cur = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(cur)
    for _ in range(reps):
        fft.exec_marked(occ, out, ws)
    b.record(cur)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    # the direct summation at the units the current field's step reads
    _abi.check(_abi.lib.svimx_nr_select(st, sim.conv.plan))
    direct = torch.full((g.size,), float("nan"), dtype=torch.float32, device=ls.device)
    _abi.check(_abi.lib.svimx_nr_sum(st, sim.conv.plan, _abi.ptr(ls.occupancy), _abi.ptr(direct), _abi.ptr(ls.ws),
                                     ls.ws.numel()))
    torch.cuda.synchronize()
    ok = ~torch.isnan(direct)
    err = float((out[ok] - direct[ok]).abs().max())
    n = int(g.size)
    print(json.dumps({"nr_fft": True, "grid": list(g.shape), "nodes": n, "ms_per_evaluation": round(ms, 4),
                      "direct_nodes_compared": int(ok.sum()), "max_abs_diff_vs_direct": err,
                      "half_spectrum_bytes_per_node_pass": 8 * (g.shape[2] // 2 + 1) / g.shape[2]}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--resolution", type=int, default=400)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-events", action="store_true", help="time without per-stage HIP events")
    ap.add_argument("--batch-jobs", type=int, default=12, help="fragment-mix jobs of the batched leg (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=12, help="requests of the served leg (0: skip)")
    ap.add_argument("--serve-load", type=float, default=2.0,
                    help="offered load of the served leg as a multiple of the measured sequential throughput")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="arrivals per second of the served leg (default: --serve-load x the sequential rate)")
    ap.add_argument("--workers", type=int, default=2, help="set-up / meshing threads of the batched forms")
    ap.add_argument("--progress", action="store_true", help="print each completed job of the batched legs (stderr)")
    ap.add_argument("--records", default=None, metavar="PATH",
                    help="append one JSON line per job of the batched and served legs to PATH")
    ap.add_argument("--batch-schedule", choices=["lockstep", "longest", "fifo", "short"], default="lockstep")
    ap.add_argument("--serve-schedule", choices=["fifo", "longest", "lockstep", "short"], default="short")
    ap.add_argument("--serve-part", default=None, metavar="served|fifo:A:B",
                    help="run one part of a long served measurement and exit (records to --records)")
    ap.add_argument("--summarize-records", nargs="+", default=None, metavar="PATH",
                    help="combine per-request records of the served parts into one comparison and exit")
    ap.add_argument("--replay-fifo", action="store_true",
                    help="served leg: run the FIFO baseline back to back and replay the queue from its service times")
    ap.add_argument("--budget-nodes", type=int, default=None, help="band nodes a tick may hold (schedule budget)")
    ap.add_argument("--slack", type=float, default=None, help="tick budget over the first job's band, as a fraction")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of this many MiB (profiler byte calibration)")
    ap.add_argument("--full-erosion", choices=["scan400", "fragment256", "cube64"], default=None,
                    help="diagnostic: run erode() to the workload's target volume end to end and exit")
    ap.add_argument("--nr-fft", action="store_true",
                    help="time the FFT N_r method on the workload and compare it with the direct sum, then exit")
    ap.add_argument("--nr-shape-sweep", action="store_true",
                    help="diagnostic: time every N_r summation shape on the workload's field "
                         "(5 launches each, in the autotuner's order) and exit")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this workload is a single erosion job on one MI355X "
              "(multi-GPU sharding is by job and is not run here); use --gpus 1", file=sys.stderr)
        return 2
    if args.summarize_records:
        return summarize_records(args.summarize_records)
    import torch
    torch.cuda.set_device(0)
    if args.serve_part:
        return served_part(args)
    if args.full_erosion:
        return full_erosion(args)
    if args.copy_calibration:
        # a plain device copy of known size: its FETCH_SIZE / WRITE_SIZE counters
        # calibrate the HBM byte counts read for the workload's kernels
        nbytes = int(args.copy_calibration) << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print(f"[bench] calibration copy of {nbytes} bytes x3 done", file=sys.stderr, flush=True)
    mesh, grid, ls, sim, law, info = build_workload(args)
    print(f"[bench] workload ready: {info}", file=sys.stderr, flush=True)
    if args.nr_shape_sweep:
        return nr_shape_sweep(sim)
    if args.nr_fft:
        return nr_fft_timing(ls, sim)
    sim.run(target_fraction=None, max_steps=args.warmup)
    torch.cuda.synchronize()
    if not args.no_events:
        # the timed steps record HIP events around every 4th N_r launch only
        # (the roofline's kernel, sampled over the whole timed region): each
        # bracketed launch costs ~5 us of device gap per event record
        sim.kernel_events = {}
        sim.kernel_event_names = {"nr_sum"}
        sim.kernel_event_every = 4
    h0 = len(sim.history)
    t0 = time.perf_counter()
    sim.run(target_fraction=None, max_steps=args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    steps = len(sim.history) - h0
    elapsed = t1 - t0
    value = steps / elapsed
    reinits = sum(1 for r in sim.history[h0:] if r.reinit)
    timed = {k: float(np.sum(v)) for k, v in sim.kernel_times_ms().items()} if not args.no_events else {}
    launches = {k: len(v) for k, v in (sim.kernel_events or {}).items()}
    sim.kernel_events = None
    sim.kernel_event_every = 1
    # N_r outputs and reads of the field the timed steps ended on
    n_out, n_in, tiles = nr_read_bytes(sim)
    stage_ms = {}
    if not args.no_events:
        # per-stage breakdown from a second, untimed pass of the same length
        sim.kernel_events = {}
        sim.kernel_event_names = None
        sim.run(target_fraction=None, max_steps=args.steps)
        stage_ms = {k: float(np.sum(v)) for k, v in sim.kernel_times_ms().items()}
    sim.kernel_events = None
    print(f"[bench] {steps} steps in {elapsed:.3f}s ({reinits} reinit); N_r ms in them {timed}; stage ms of an "
          f"untimed pass {stage_ms}", file=sys.stderr, flush=True)
    # roofline of the dominant kernel (the N_r summation)
    nr_ms = timed.get("nr_sum", 0.0) / max(launches.get("nr_sum", 1), 1)
    alg_bytes = 4.0 * (n_out + n_in)
    # FP32 rate in ball taps (2 flops per nonzero weight and output, the dense
    # convolution's count); the kernels execute fewer FMAs than that, since
    # the +dx and -dx windows share their columns' taps
    taps = float((sim.conv.weights > 0).sum())
    flops = 2.0 * taps * n_out
    achieved = alg_bytes / (nr_ms * 1e-3) / 1e9 if nr_ms > 0 else 0.0
    conv = sim.conv
    if conv.method == "fft":
        kname = "k_fft_lines"
    elif conv.run_length == 0:
        kname = "k_nr_planes"
    else:
        kname = "k_nr_lds" if conv.staging == "lds" else ("k_nr_lines" if conv.lines > 1 else "k_nr_runs")
    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"{kname}_traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if t.get("shape", conv.shape_name) == conv.shape_name:
            traffic = t["read_bytes_per_dispatch"] + t["write_bytes_per_dispatch"]
    # The N_r summation does ~2.9k flops per output on ~17 bytes per output of
    # algorithmic traffic (~170 flop/B, far right of the ~20 flop/B ridge): its
    # roofline is the FP32 vector peak.  achieved = the dense convolution's flops
    # (2 per nonzero ball weight and output) / the live launch time; the HBM
    # figures stay alongside.  Measured limiter (profiles/): the vector-memory
    # address path (one cache line per lane of its window loads), not VALU
    # issue or HBM.
    tflops = flops / (nr_ms * 1e-3) / 1e12 if nr_ms > 0 else 0.0
    roofline = {"bound": "valu", "achieved": round(tflops, 2), "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tflops / FP32_VALU_PEAK_TFLOPS, 4), "traffic": traffic,
                "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated (profiles/)" if traffic else None,
                "kernel": kname, "avg_ms": round(nr_ms, 4), "flops_per_launch": flops,
                "frac_unpacked_fp32": round(tflops / (FP32_VALU_PEAK_TFLOPS / 2), 4),
                "alg_bytes_per_launch": alg_bytes, "hbm_achieved_gbs": round(achieved, 1),
                "hbm_frac": round(achieved / HBM_PEAK_GBS, 4),
                "outputs_per_launch": n_out, "tiles": tiles}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(ls, law)
    del sim, ls
    torch.cuda.empty_cache()
    batched = served = None
    seq_rate = None
    from svierosion_amd.batch import DEFAULT_BUDGET_NODES, DEFAULT_SLACK
    budget = DEFAULT_BUDGET_NODES if args.budget_nodes is None else args.budget_nodes
    slack = DEFAULT_SLACK if args.slack is None else args.slack
    if args.batch_jobs > 0:
        batched, rs = batched_leg(args.batch_jobs, args.workers, args.batch_schedule, budget, slack, args.progress,
                                  args.records)
        seq_rate = len(rs.done) / rs.makespan if rs.makespan > 0 else None
    if args.serve_requests > 0:
        rate = args.serve_rate or args.serve_load * (seq_rate if seq_rate else 0.5)
        served = served_leg(args.serve_requests, rate, args.workers, args.serve_schedule, budget, slack,
                            args.progress, args.records, args.replay_fifo)
    line = {"metric": "RK3 steps/s + HBM GB/s @res400 1.3M-tri; batched erosions/hour, p99 latency",
            "value": round(value, 3), "unit": "RK3 steps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": round(1000 * elapsed / max(steps, 1), 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "1.3M-triangle noisy ellipsoid scan, SVI k=2 s=1 t=0.95 r=2.0, res 400",
                       "triangles": info["triangles"], "grid": info["grid"], "band_nodes": info["band_nodes"],
                       "reinit_in_window": reinits,
                       "stage_ms_untimed_pass": {k: round(v, 3) for k, v in stage_ms.items()},
                       "nr_shape": info["nr_shape"], "nr_tuning_ms": info["nr_tuning_ms"],
                       "launches": launches},
            "roofline": roofline, "cpu_baseline": cpu, "batched": batched, "served": served}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
