#!/usr/bin/env python3
"""Agent-updates/s of the flockers agent step on one MI355X.

Workload (the plain GPU drop-in): 1,048,576 boids on a 20,480 x 20,480
toroidal Field2D, discretization 10/1.5, neighbour radius 10, all weights 1,
jump 0.7, positions uniform and headings unit vectors at a uniform angle
from the seeded Philox stream.  One step = one Schedule::step of the
flockers model: neighbour query + Bird::step for every agent and the
rebinning of the field (lazy_update).  State is resident in HBM before the
timed region starts.

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
# Algorithmic bytes per agent and launch (f32 x, y, hx, hy + u32 id = 20 B):
BYTES_RT = 20 + 20       # tile-resident step: read the agent, write it back
BYTES_FLOCK = 20 + 16    # three-kernel path, fused query + step
BYTES_SCATTER = 20 + 20  # three-kernel path, rebinning
BYTES_STEP = 76          # per agent-update of a whole step (binning 40 + query/step 36)


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

    def elapsed_ms(self, a, b):
        self._ok(self.L.hipEventSynchronize(b), "hipEventSynchronize")
        ms = ctypes.c_float()
        self._ok(self.L.hipEventElapsedTime(ctypes.byref(ms), a, b), "hipEventElapsedTime")
        return ms.value


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def plain_workload():
    return dict(n=1 << 20, side=20480.0, d=10.0 / 1.5, radius=10.0, seed=1)


def timed_run(hip, f, p, first, steps, binned):
    e0, e1 = hip.event(), hip.event()
    hip.sync()
    t0 = time.perf_counter()
    hip.record(e0, f.stream)
    f.schedule_run(p, first, steps, binned=binned)
    hip.record(e1, f.stream)
    ms = hip.elapsed_ms(e0, e1)
    hip.sync()
    return ms, time.perf_counter() - t0


def run_plain(args, hip):
    from krabmaga_amd import Field2D, FlockParams
    w = plain_workload()
    n, side = w["n"], w["side"]
    p = FlockParams(radius=w["radius"], seed=w["seed"])
    f = Field2D(side, side, w["d"], True, n)
    f.init_uniform(n, w["seed"], 1)
    f.schedule_run(p, 0, args.warmup)  # step 0 includes the first lazy_update
    f.sync()
    log(f"plain warmup done ({args.warmup} steps)")
    ms, wall = timed_run(hip, f, p, args.warmup, args.steps, binned=False)
    f.sync()
    layout = f.stats()
    log(f"tile-resident timed region: {ms:.3f} ms for {args.steps} steps; layout {layout}")

    # per-kernel durations: steps launched one by one, each launch bracketed
    # by HIP events on the field's stream
    first = args.warmup + args.steps
    kprof = min(args.steps, args.profile_steps)
    f.profile(True)
    f.schedule_run(p, first, kprof, use_graph=False)
    rt_ms, rt_n = f.profile_read("rt_step")
    f.profile(False)
    first += kprof

    # the three-kernel path on the same state, for comparison
    nb = min(args.steps, args.binned_steps)
    bms, _ = timed_run(hip, f, p, first, nb, binned=True)
    first += nb
    f.profile(True)
    f.schedule_run(p, first, kprof, use_graph=False, binned=True)
    kern = {"rt_step": {"avg_us": 1e3 * rt_ms / max(rt_n, 1), "launches": rt_n}}
    for k in ("flock", "scan", "scatter"):
        tot, cnt = f.profile_read(k)
        kern[k] = {"avg_us": 1e3 * tot / max(cnt, 1), "launches": cnt}
    f.profile(False)
    f.close()
    log(f"three-kernel path: {bms / nb:.4f} ms/step")
    return w, ms, wall, kern, bms / nb, layout


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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--binned-steps", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    hip = Hip()
    w, ms, wall, kern, binned_ms, layout = run_plain(args, hip)
    n, steps = w["n"], args.steps
    value = n * steps / (ms * 1e-3)
    # dominant kernel of the timed path: the tile-resident step, or the
    # fused flock kernel when the field ran the three-kernel path
    if kern["rt_step"]["launches"]:
        dom, dom_bytes = "rt_step", BYTES_RT
    else:
        dom, dom_bytes = "flock", BYTES_FLOCK
    dom_us = kern[dom]["avg_us"]
    achieved = dom_bytes * n / (dom_us * 1e-6) if dom_us > 0 else 0.0
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
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_" + dom,
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": None,
            "bytes_per_agent": dom_bytes,
        },
        "step_roofline": {
            "bytes_per_agent_update": BYTES_STEP,
            "achieved_GBps": BYTES_STEP * value / 1e9,
            "frac": BYTES_STEP * value / HBM_PEAK,
        },
        "kernels_us": {k: round(v["avg_us"], 3) for k, v in kern.items()},
        "three_kernel_path": {
            "ms_per_step": binned_ms,
            "agent_updates_per_s": n / (binned_ms * 1e-3),
        },
        "wall_s_timed_region": wall,
        "layout": {k: layout[k] for k in ("rt_active", "rt_plans", "rt_flags", "tiles",
                                          "dense_tiles", "lds_bytes", "rt_steps")},
    }
    if not args.no_cpu:
        log("CPU baseline (oracle, 1 thread) ...")
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
