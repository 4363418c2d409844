"""Serving loop: exploration runs arrive over time and share the GPU.

A request is a flockers run (RunSpec) and a number of schedule steps.
Requests are admitted in arrival order while their device footprint fits the
memory budget; admitted runs live in one Ensemble, so every tick advances all
of them with one launch per schedule step (per tile class).  A tick lasts
until the earliest active run completes (at most `max_tick` steps), so each
run stops exactly at its own step count; completed runs are read back and
retired, which frees their memory for the queue.  Requests that arrive while
others are on the GPU are admitted at the next tick (continuous batching),
and a run's result is the bits it would produce alone.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from .ensemble import Ensemble, footprint


@dataclass
class Completed:
    index: int          # position in the request list
    arrival: float      # seconds since the stream started
    admitted: float
    done: float
    agents: int
    steps: int
    state: dict = field(default=None, repr=False)

    @property
    def latency(self):
        return self.done - self.arrival


class Server:
    def __init__(self, hbm_budget, tuning=None, max_tick=64, keep_state=False, clock=None,
                 ensemble=None, footprint_fn=None, admit_every=1):
        self.budget = int(hbm_budget)
        self.tuning = tuning
        self.max_tick = int(max_tick)
        self.keep_state = keep_state
        self.clock = clock or time.perf_counter
        self.make_ensemble = ensemble or (lambda: Ensemble(self.tuning))
        self.footprint = footprint_fn or (lambda spec: footprint(spec, self.tuning))
        # while runs are active, new ones join at most every `admit_every`
        # ticks: each admission rebuilds the layout
        self.admit_every = max(1, int(admit_every))
        self.ticks = 0
        self.admissions = 0
        self.peak_bytes = 0

    def serve(self, requests, progress=None):
        """Runs the request stream to completion against the wall clock;
        returns the Completed records in completion order.  progress(msg),
        if given, is called every 16 ticks."""
        reqs = sorted(enumerate(requests), key=lambda r: r[1].arrival)
        fp = [self.footprint(r.spec) for _, r in reqs]
        for (i, r), b in zip(reqs, fp):
            if b > self.budget:
                raise ValueError(f"request {i} needs {b} bytes, more than the budget")
        done = []
        active = {}  # run id -> [request index, request, remaining steps, admit time, bytes]
        used = 0
        head = 0
        ens = self.make_ensemble()
        t0 = self.clock()
        last_admit = -self.admit_every
        try:
            while head < len(reqs) or active:
                now = self.clock() - t0
                batch = []
                may_admit = not active or self.ticks - last_admit >= self.admit_every
                while may_admit and head < len(reqs) and reqs[head][1].arrival <= now and \
                        used + fp[head] <= self.budget:
                    batch.append((reqs[head], fp[head]))
                    used += fp[head]
                    head += 1
                self.peak_bytes = max(self.peak_bytes, used)
                if batch:
                    ids = ens.admit([r.spec for (_, r), _ in batch])
                    self.admissions += 1
                    last_admit = self.ticks
                    for rid, ((i, r), b) in zip(ids, batch):
                        active[rid] = [i, r, r.steps, now, b]
                if not active:
                    # idle until the next arrival
                    wait = reqs[head][1].arrival - (self.clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                    continue
                tick = min(self.max_tick, min(a[2] for a in active.values()))
                ens.run(tick)
                ens.sync()
                self.ticks += 1
                if progress and self.ticks % 16 == 0:
                    progress(f"tick {self.ticks}: {len(active)} active, {len(done)} done, "
                             f"{len(reqs) - head} queued, {self.clock() - t0:.1f} s")
                for rid in [k for k, a in active.items() if a[2] == tick]:
                    i, r, _, adm, b = active.pop(rid)
                    state = ens.read(rid) if self.keep_state else None
                    ens.retire(rid)
                    used -= b
                    # a request is done once its result is on the host
                    done.append(Completed(i, r.arrival, adm, self.clock() - t0, r.spec.agents,
                                          r.steps, state))
                for a in active.values():
                    a[2] -= tick
            self.wall = self.clock() - t0
            self.layout = ens.stats() if hasattr(ens, "stats") else {}
        finally:
            ens.close()
        return done


def summary(done, wall):
    lat = np.array([c.latency for c in done])
    upd = float(sum(c.agents * c.steps for c in done))
    return {
        "requests": len(done),
        "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "mean_latency_s": float(lat.mean()) if len(lat) else None,
        "agent_updates_per_s": upd / wall if wall > 0 else None,
        "wall_s": wall,
    }
