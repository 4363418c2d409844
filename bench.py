#!/usr/bin/env python3
"""Agent-updates/s of the flockers agent step on one MI355X.

Workload (the plain GPU drop-in): 1,048,576 boids on a 20,480 x 20,480
toroidal Field2D, discretization 10/1.5, neighbour radius 10, all weights 1,
jump 0.7, positions uniform and headings unit vectors at a uniform angle
from the seeded Philox stream.  One step = one Schedule::step of the
flockers model: neighbour query + Bird::step for every agent and the
rebinning of the field (lazy_update).  State is resident in HBM before the
timed region starts.  The headline runs the field's default path (at this
size the three-kernel path: k_flock, then k_rebin_gather, which also does
the bin-count scan, then k_rebin_fallback, the scatter and bin sort that
run only in a step the gather could not rebin); the tile-resident step
(one k_rt_step launch per tile class) is timed beside it from the same
initial state.

Beside it, in the same JSON line: the clustered-density workload (8,388,608
boids, 64 Gaussian blobs), the batched ensemble (512 replicas of uneven
size, one launch per step for all of them, layout chosen by the autotuner)
against the same runs launched one by one, and the serving loop (Poisson
arrivals admitted by HBM footprint) with p50/p99 latency, a sweep of the
arrival rate and the oracle's worker pool on the host cores.

Prints one JSON line.  `python bench.py --gpus N` with N > 1 exits: this
build measures one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "agent-updates/s (1M flockers f32) & ensemble p50/p99 latency, batched vs plain"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
# Algorithmic bytes per agent-update of a whole schedule step (f32 x, y, hx,
# hy + u32 id = 20 B per agent): binning reads and writes the agent (40 B),
# the fused query + Bird::step reads it and writes position and heading
# (36 B).  The roofline prices every kernel split of the step at this figure.
BYTES_STEP = 76
# What individual kernels move per agent (reported beside it):
BYTES_RT = 20 + 20       # tile-resident step: read the agent, write it back
BYTES_FLOCK = 20 + 16    # three-kernel path, fused query + step
CPU_SHARE = 16           # host cores that go with one GPU of the box


class Hip:
    """The few HIP runtime calls the timing needs."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def sync(self):
        self._ok(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def event(self):
        e = ctypes.c_void_p()
        self._ok(self.L.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e

    def record(self, e, stream):
        self._ok(self.L.hipEventRecord(e, ctypes.c_void_p(stream)), "hipEventRecord")

    def copy_calibration(self, nbytes):
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        self._ok(self.L.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)), "hipMalloc")
        self._ok(self.L.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)), "hipMalloc")
        self._ok(self.L.hipMemset(a, 1, ctypes.c_size_t(nbytes)), "hipMemset")
        self._ok(self.L.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3), "hipMemcpy")
        self.sync()
        self.L.hipFree(a)
        self.L.hipFree(b)

    def elapsed_ms(self, a, b):
        self._ok(self.L.hipEventSynchronize(b), "hipEventSynchronize")
        ms = ctypes.c_float()
        self._ok(self.L.hipEventElapsedTime(ctypes.byref(ms), a, b), "hipEventElapsedTime")
        return ms.value


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores():
    """Cores this process may run on, and the threads a CPU baseline uses:
    the box's share for one GPU (16), or fewer if fewer are available."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return avail, max(1, min(CPU_SHARE, avail))


def plain_workload():
    return dict(n=1 << 20, side=20480.0, d=10.0 / 1.5, radius=10.0, seed=1)


def timed_run(hip, f, p, first, steps, binned=False, tile_resident=False):
    e0, e1 = hip.event(), hip.event()
    hip.sync()
    t0 = time.perf_counter()
    hip.record(e0, f.stream)
    f.schedule_run(p, first, steps, binned=binned, tile_resident=tile_resident)
    hip.record(e1, f.stream)
    ms = hip.elapsed_ms(e0, e1)
    hip.sync()
    return ms, time.perf_counter() - t0


def profiled_traffic(name):
    """HBM bytes per schedule step (all kernels of the step), from a committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a 1 GiB
    copy), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_step:"):
                return int(line.split(":")[1]), "profiles/" + name
    except OSError:
        pass
    return None, None


PLAIN_KERNELS = {"default": ("flock", "scan", "scatter", "sort_bins"),
                 "tile_resident": ("rt_step",)}


def run_plain(args, hip):
    """The 1M-boid field on its default path (the three-kernel path at this
    size) and, from the same initial state, on the tile-resident step: each
    warmed up, then timed over the same window.  Per-kernel durations come
    from the same run made again from the same initial state with the timed
    window's steps launched one by one (HIP events around every launch); the
    run is deterministic, so they cover the very steps the timed region ran."""
    from krabmaga_amd import Field2D, FlockParams
    w = plain_workload()
    n, side = w["n"], w["side"]
    p = FlockParams(radius=w["radius"], seed=w["seed"])
    out = {}
    for path in ("default", "tile_resident"):
        kw = {"tile_resident": path == "tile_resident"}
        f = Field2D(side, side, w["d"], True, n)
        f.init_uniform(n, w["seed"], 1)
        f.schedule_run(p, 0, args.warmup, **kw)  # step 0 includes the first lazy_update
        f.sync()
        ms, wall = timed_run(hip, f, p, args.warmup, args.steps, **kw)
        f.sync()
        layout = f.stats()
        log(f"{path} path: {ms / args.steps:.4f} ms/step over {args.steps} steps; "
            f"layout {layout}")
        f.close()
        kprof = args.steps if args.profile_steps <= 0 else min(args.steps, args.profile_steps)
        f = Field2D(side, side, w["d"], True, n)
        f.init_uniform(n, w["seed"], 1)
        f.schedule_run(p, 0, args.warmup, **kw)
        f.sync()
        f.profile(True)
        f.schedule_run(p, args.warmup, kprof, use_graph=False, **kw)
        kern = {}
        for k in PLAIN_KERNELS[path]:
            tot, cnt = f.profile_read(k)
            kern[k] = {"avg_us": 1e3 * tot / max(cnt, 1), "launches": cnt,
                       "us_per_step": 1e3 * tot / kprof}
        f.profile(False)
        f.close()
        out[path] = {"ms": ms, "wall": wall, "layout": layout, "kern": kern}
    return w, out


def cpu_baseline(args):
    """The oracle (single thread) on a bounded sample of the same workload."""
    from oracle import oracle as O
    w = plain_workload()
    n, side = w["n"], w["side"]
    x, y, hx, hy = O.init_uniform(n, side, side, w["seed"], 1)
    secs, _ = O.time_steps(side, side, w["d"], True, x, y, hx, hy,
                           O.params(radius=w["radius"], seed=w["seed"]), args.cpu_steps)
    return {
        "value": n * args.cpu_steps / secs,
        "unit": "agent-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{args.cpu_steps} schedule steps of the same 1,048,576-boid run "
                  f"(oracle restatement of Field2D + Bird::step, {secs:.1f} s)",
    }


def _timed(hip, stream, fn):
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    fn()
    hip.record(e1, stream)
    return hip.elapsed_ms(e0, e1)


def _one_by_one(hip, jobs, warmup, whole_run=False):
    """The same runs launched one by one with the plain kernels, a field per
    run.  whole_run=False: each run warmed up, then its steps timed with HIP
    events (state resident, as the batched figure); True: the wall time of
    each run's whole life -- field, Flocker::init, every step, read-back --
    as a served request has it.  Returns the summed milliseconds."""
    from krabmaga_amd import Field2D
    total_ms = 0.0
    for spec, steps in jobs:
        if whole_run:
            t0 = time.perf_counter()
            f = Field2D(spec.width, spec.height, spec.discretization, True, spec.agents)
            f.init_uniform(spec.agents, spec.init_seed, spec.heading_mode)
            f.schedule_run(spec.params, 0, steps)
            f.read_agents(0)
            total_ms += 1e3 * (time.perf_counter() - t0)
            f.close()
            continue
        f = Field2D(spec.width, spec.height, spec.discretization, True, spec.agents)
        f.init_uniform(spec.agents, spec.init_seed, spec.heading_mode)
        f.schedule_run(spec.params, 0, warmup)
        f.sync()
        total_ms += _timed(hip, f.stream, lambda: f.schedule_run(spec.params, warmup, steps))
        f.close()
    return total_ms


def run_batched(args, hip):
    """512 replicas of uneven size in one layout, its form and tile shape
    chosen by the autotuner."""
    from krabmaga_amd import Ensemble, autotune, workloads
    specs = workloads.batched(args.replicas)
    agents = sum(s.agents for s in specs)
    steps = args.batched_steps
    log(f"batched: {len(specs)} replicas, {agents} agents")
    table = []
    tuning, path = None, {}
    if not args.no_tune:
        tuning, path, table = autotune.tune(specs, steps=16, warmup=8, log=log)
    ens = Ensemble(tuning)
    ens.admit(specs)
    ens.run(args.warmup, **path)
    ens.sync()
    ms = _timed(hip, ens.stream, lambda: ens.run(steps, **path))
    st = ens.stats()
    ens.close()
    value = agents * steps / (ms * 1e-3)
    log(f"batched: {ms / steps:.3f} ms/step, {value:.3e} agent-updates/s, layout {st}")
    # one by one with the plain kernels, on a bounded sample of the runs
    k = min(len(specs), args.one_by_one_runs)
    sample = specs[:: max(1, len(specs) // k)][:k]
    sample_agents = sum(s.agents for s in sample)
    obo_ms = _one_by_one(hip, [(s, steps) for s in sample], args.warmup)
    obo_value = sample_agents * steps / (obo_ms * 1e-3)
    log(f"one by one: {len(sample)} runs, {obo_value:.3e} agent-updates/s")
    out = {
        "workload": f"{len(specs)} flockers replicas, agents log-uniform in [1000, 262144], "
                    f"constant density 0.025, weights U[0.5, 2], jump U[0.3, 1], {steps} steps",
        "replicas": len(specs), "agents": agents, "steps": steps,
        "value": value, "unit": "agent-updates/s", "ms_per_step": ms / steps,
        "one_by_one": {"value": obo_value, "runs": len(sample), "agents": sample_agents,
                       "sample": "every %d-th replica, each on its own field" %
                                 max(1, len(specs) // k)},
        "vs_one_by_one": value / obo_value,
        "tuning": None if tuning is None else tuning.__dict__,
        "form": "bin-grouped (three-kernel path)" if path.get("binned") else
                ("tile-resident" if path.get("tile_resident") else "default"),
        "autotune": [{"tuning": t.__dict__, "path": pth, "ms_per_step": m} for t, pth, m in table],
        "layout": st,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_ensemble(specs, args.cpu_threads)
    return out


def cpu_ensemble(specs, threads):
    """explore_parallel!'s rayon path: the oracle on a thread pool, one run
    per thread at a time, bounded to one schedule step of every replica."""
    from oracle import oracle as O
    avail, share = host_cores()
    nthreads = threads or share
    jobs = [{"n": s.agents, "side": s.width, "discretization": s.discretization, "steps": 1,
             "params": O.params(cohesion=s.params.cohesion, avoidance=s.params.avoidance,
                                consistency=s.params.consistency,
                                randomness=s.params.randomness, momentum=s.params.momentum,
                                jump=s.params.jump, radius=s.params.radius, seed=s.params.seed),
             "heading_mode": s.heading_mode} for s in specs]
    wall, _, _ = O.run_ensemble(jobs, nthreads)
    agents = sum(s.agents for s in specs)
    return {"value": agents / wall, "unit": "agent-updates/s", "cores": nthreads,
            "host_cores": avail, "kind": "port",
            "sample": f"1 schedule step of each of the {len(specs)} replicas "
                      f"({agents} agent-updates, {wall:.1f} s) on a pool of {nthreads} threads "
                      f"(the box's {CPU_SHARE}-core share per GPU; {avail} cores visible)"}


def run_clustered(args, hip):
    """8,388,608 boids on a 40,960 x 40,960 torus, 80 % in 64 Gaussian blobs
    (sigma 200) and 20 % uniform: a bounded number of timed steps."""
    import numpy as np
    from krabmaga_amd import Field2D, FlockParams, workloads
    n, side, seed = 8_388_608, 40_960.0, 2
    t0 = time.perf_counter()
    x, y, hx, hy = workloads.clustered(n, side, blobs=64, sigma=200.0, uniform_frac=0.2,
                                       seed=seed)
    log(f"clustered: {n} boids generated in {time.perf_counter() - t0:.1f} s")
    p = FlockParams(seed=seed)
    f = Field2D(side, side, 10.0 / 1.5, True, n)
    f.set_object_locations(np.arange(n, dtype=np.uint32), x, y, hx, hy)
    del x, y, hx, hy
    warm = args.clustered_warmup
    f.schedule_run(p, 0, warm)
    f.sync()
    steps = args.clustered_steps
    ms, _ = timed_run(hip, f, p, warm, steps)
    st = f.stats()
    # per-kernel time over a few steps launched one by one
    f.profile(True)
    f.schedule_run(p, warm + steps, args.clustered_profile_steps, use_graph=False)
    kern = {}
    for k in ("rt_step", "flock", "scan", "scatter", "sort_bins"):
        tot, cnt = f.profile_read(k)
        if cnt:
            kern[k] = {"total_us_per_step": 1e3 * tot / args.clustered_profile_steps,
                       "launches": cnt}
    f.profile(False)
    f.close()
    value = n * steps / (ms * 1e-3)
    log(f"clustered: {ms / steps:.3f} ms/step, {value:.3e} agent-updates/s")
    return {
        "workload": "flockers clustered density: 8,388,608 boids, 40,960 x 40,960 torus, 64 "
                    "Gaussian blobs (sigma 200) + 20 % uniform, radius 10, weights 1, jump 0.7",
        "agents": n, "warmup": warm, "steps": steps, "ms_per_step": ms / steps,
        "value": value, "unit": "agent-updates/s",
        "path": "tile-resident" if st["rt_steps"] else "three-kernel (flock + scan + scatter "
                "+ bin sort)",
        "roofline": {"bytes_per_agent_update": BYTES_STEP,
                     "achieved_GBps": BYTES_STEP * value / 1e9,
                     "frac": BYTES_STEP * value / HBM_PEAK},
        "kernels_us_per_step": {k: round(v["total_us_per_step"], 1) for k, v in kern.items()},
        "layout": {k: st[k] for k in ("rt_active", "rt_steps", "tiles", "dense_tiles")},
    }


def cpu_serving(reqs, duration, max_steps):
    """The rayon path for the serving stream: the oracle on a worker pool,
    one request per thread at a time, each request bounded to `max_steps`
    schedule steps."""
    from oracle import oracle as O
    avail, nthreads = host_cores()
    jobs = []
    for r in reqs:
        s, pp = r.spec, r.spec.params
        jobs.append({"n": s.agents, "side": s.width, "discretization": s.discretization,
                     "steps": min(r.steps, max_steps),
                     "params": O.params(cohesion=pp.cohesion, avoidance=pp.avoidance,
                                        consistency=pp.consistency, randomness=pp.randomness,
                                        momentum=pp.momentum, jump=pp.jump, radius=pp.radius,
                                        seed=pp.seed),
                     "heading_mode": s.heading_mode})
    wall, per_run, _ = O.run_ensemble(jobs, nthreads)
    upd = sum(j["n"] * j["steps"] for j in jobs)
    value = upd / wall
    work = sum(r.spec.agents * r.steps for r in reqs) / len(reqs)
    return {"value": value, "unit": "agent-updates/s", "cores": nthreads, "host_cores": avail,
            "kind": "port",
            "sustainable_rate_runs_per_s": value / work,
            "sample": f"the {len(reqs)} requests of the stream, each run for min(steps, "
                      f"{max_steps}) schedule steps on a pool of {nthreads} threads ({upd} "
                      f"agent-updates, {wall:.1f} s); a request's full run averages "
                      f"{work:.3g} agent-updates, so the pool sustains "
                      f"{value / work:.3g} runs/s against {len(reqs) / duration:g} offered"}


def serving_sweep(args):
    """The serving loop at several Poisson arrival rates: served against
    offered load and the latency at each."""
    from krabmaga_amd import serve, workloads
    out = []
    for rate in args.sweep_rates:
        dur = max(args.sweep_min_seconds, args.sweep_requests / rate)
        reqs = workloads.serving(rate=rate, duration=dur, seed=40 + int(rate))
        if not reqs:
            continue
        offered = sum(r.spec.agents * r.steps for r in reqs) / dur
        srv = serve.Server(args.serve_budget_gb * 2**30, keep_state=True)
        done = srv.serve(reqs)
        sm = serve.summary(done, srv.wall)
        row = {"rate_runs_per_s": rate, "requests": len(reqs), "arrivals_s": dur,
               "offered_agent_updates_per_s": offered,
               "served_agent_updates_per_s": sm["agent_updates_per_s"],
               "p50_latency_s": sm["p50_latency_s"], "p99_latency_s": sm["p99_latency_s"],
               "wall_s": sm["wall_s"]}
        log(f"serving sweep: {row}")
        out.append(row)
    return out


def run_serving(args, hip):
    """Poisson arrivals of exploration runs, admitted by HBM footprint."""
    from krabmaga_amd import serve, workloads
    reqs = workloads.serving(rate=args.serve_rate, duration=args.serve_seconds)
    offered = sum(r.spec.agents * r.steps for r in reqs) / args.serve_seconds
    log(f"serving: {len(reqs)} requests over {args.serve_seconds} s")
    srv = serve.Server(args.serve_budget_gb * 2**30, keep_state=True)
    done = srv.serve(reqs, progress=log)
    sm = serve.summary(done, srv.wall)
    log(f"serving: {sm}")
    k = min(len(reqs), args.one_by_one_runs)
    # k requests spread evenly over the arrival order
    m = min(k, len(reqs))
    idx = sorted({round(i * (len(reqs) - 1) / max(m - 1, 1)) for i in range(m)})
    sample = [reqs[i] for i in idx]
    obo_ms = _one_by_one(hip, [(r.spec, r.steps) for r in sample], args.warmup, whole_run=True)
    obo_value = sum(r.spec.agents * r.steps for r in sample) / (obo_ms * 1e-3)
    log(f"serving one by one: {len(sample)} runs, {obo_value:.3e} agent-updates/s")
    sm.update({
        "workload": f"Poisson arrivals at {args.serve_rate:g} runs/s for {args.serve_seconds:g} s, "
                    "agents log-uniform in [1000, 1e6], steps uniform in [50, 2000], weights "
                    "U[0.5, 2], jump U[0.3, 1]",
        "offered_agent_updates_per_s": offered,
        "hbm_budget_bytes": srv.budget, "peak_admitted_bytes": srv.peak_bytes,
        "ticks": srv.ticks, "admissions": srv.admissions, "layout": srv.layout,
        "one_by_one": {"value": obo_value, "runs": len(sample),
                       "sample": "%d of the %d requests, evenly spaced in arrival order, each "
                                 "run alone on its own field: creation, init, all steps and "
                                 "read-back timed, as a served request" % (len(sample), len(reqs))},
        "vs_one_by_one": sm["agent_updates_per_s"] / obo_value,
    })
    if args.sweep_rates:
        sweep = serving_sweep(args)
        # the highest rate whose offered load was served (its stream drained
        # within 1.25x its arrival window)
        ok = [r for r in sweep if r["wall_s"] <= 1.25 * r["arrivals_s"]]
        sm["rate_sweep"] = sweep
        sm["sustainable"] = max(ok, key=lambda r: r["rate_runs_per_s"]) if ok else None
    if not args.no_cpu:
        log("serving CPU baseline (oracle worker pool) ...")
        sm["cpu_baseline"] = cpu_serving(reqs, args.serve_seconds, args.serve_cpu_steps)
    return sm


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="steps of the timed window given per-kernel durations (0: all)")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--batched-steps", type=int, default=300)
    ap.add_argument("--one-by-one-runs", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-seconds", type=float, default=0.5)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-cpu-steps", type=int, default=2)
    ap.add_argument("--sweep-rates", type=float, nargs="*", default=[5.0, 10.0, 20.0, 40.0])
    ap.add_argument("--sweep-requests", type=int, default=24)
    ap.add_argument("--sweep-min-seconds", type=float, default=1.0)
    ap.add_argument("--no-clustered", action="store_true")
    ap.add_argument("--clustered-warmup", type=int, default=3)
    ap.add_argument("--clustered-steps", type=int, default=20)
    ap.add_argument("--clustered-profile-steps", type=int, default=4)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also copy 1 GiB device to device (a known-size reference for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the engine "
              "runs every field and ensemble on one GPU (no sharding across GPUs)",
              file=sys.stderr)
        sys.exit(2)

    hip = Hip()
    w, plain = run_plain(args, hip)
    if args.calibrate_copy:
        hip.copy_calibration(1 << 30)
    n, steps = w["n"], args.steps
    d, rt = plain["default"], plain["tile_resident"]
    ms = d["ms"]
    value = n * steps / (ms * 1e-3)
    three = d["layout"]["rt_steps"] == 0
    # The roofline prices the whole schedule step at the 76 algorithmic bytes
    # per agent-update: a "launch" is one schedule step -- on the default
    # (three-kernel) path its kernels back to back on the field's stream --
    # and its duration is the timed region's time per step (HIP events on
    # that stream).  The dominant kernel's own figure goes beside.
    step_us = 1e3 * ms / steps
    achieved = BYTES_STEP * n / (step_us * 1e-6)
    dom, dom_bytes = ("flock", BYTES_FLOCK) if three else ("rt_step", BYTES_RT)
    dom_us = d["kern"][dom]["avg_us"]
    prof = "flockers_1M_step.md" if three else "k_rt_step_flockers_1M.md"
    traffic, traffic_src = profiled_traffic(prof)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "agent-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Philox-seeded uniform positions, unit headings at uniform angle",
        "config": {
            "workload": "flockers plain GPU drop-in: 1,048,576 boids, 20,480 x 20,480 torus, "
                        "discretization 10/1.5, radius 10, weights 1, jump 0.7",
            "agents": n, "side": w["side"], "seed": w["seed"],
            "path": "three-kernel (k_flock, scan, k_rebin_gather), the field's default at "
                    "this size" if three else "tile-resident (k_rt_step)",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("schedule step: k_flock + k_rebin_gather (which does the bin-count "
                       "scan in this run) + k_rebin_fallback (scatter and bin sort; returns "
                       "at once unless the gather could not rebin the step)"
                       if three else "schedule step: all k_rt_step launches of the step"),
            "launch_us": step_us,
            "unit_of_launch": "one schedule step",
            "bytes_per_agent": BYTES_STEP,
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "traffic_source": traffic_src,
            # the flock phase: k_flock, plus its k_flock_global launch (tiles
            # that do not fit LDS) on fields that need one, timed as one
            "dominant_kernel": "k_flock (+ k_flock_global)" if three else "k_rt_step",
            "dominant_kernel_us": dom_us,
            "dominant_kernel_bytes_per_agent": dom_bytes,
            "dominant_kernel_frac": dom_bytes * n / (dom_us * 1e-6) / HBM_PEAK if dom_us else 0.0,
            "limiter": "issue and latency: the step moves its bytes once, but the neighbour "
                       "walk and Bird::step are VALU work per bird, and every kernel runs short "
                       "dependent chains of global and LDS round trips",
        },
        "kernels_us": {k: round(v["avg_us"], 3) for k, v in d["kern"].items()},
        "kernels_us_per_step": {k: round(v["us_per_step"], 3) for k, v in d["kern"].items()},
        "tile_resident_path": {
            "ms_per_step": rt["ms"] / steps,
            "agent_updates_per_s": n * steps / (rt["ms"] * 1e-3),
            "k_rt_step_us_per_step": round(rt["kern"]["rt_step"]["us_per_step"], 3),
            "layout": {k: rt["layout"][k] for k in ("rt_active", "rt_plans", "rt_flags", "tiles",
                                                    "dense_tiles", "lds_bytes", "rt_steps")},
        },
        "wall_s_timed_region": d["wall"],
        "layout": {k: d["layout"][k] for k in ("tiles", "tile_cols", "tile_rows", "bin_rows",
                                               "rt_steps")},
    }
    if not args.no_cpu:
        log("CPU baseline (oracle, 1 thread) ...")
        out["cpu_baseline"] = cpu_baseline(args)
    if not args.no_clustered:
        out["clustered"] = run_clustered(args, hip)
    if not args.no_batched:
        out["batched"] = run_batched(args, hip)
    if not args.no_serving:
        out["serving"] = run_serving(args, hip)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
