"""Synthetic flockers workloads: the batched ensemble and the serving stream.

Both are the shapes the explore module produces (model_exploration.rs:
354-432, parameters sampled per run): runs of uneven sizes at constant
density, with weights and jump drawn per run.  Sampling is host-side and
seeded, so a workload is a pure function of its seed.
"""
from dataclasses import dataclass

import numpy as np

from .ensemble import RunSpec
from .field import FlockParams

D = 10.0 / 1.5
DENSITY = 0.025  # birds per unit^2 of the 1,000-boid 200 x 200 flockers run


def _run(rng, n, tag, i, heading_mode=1):
    side = float(np.sqrt(n / DENSITY))
    w = rng.uniform(0.5, 2.0, size=5)
    jump = float(rng.uniform(0.3, 1.0))
    p = FlockParams(cohesion=float(w[0]), avoidance=float(w[1]), consistency=float(w[2]),
                    randomness=float(w[3]), momentum=float(w[4]), jump=jump, radius=10.0,
                    seed=(tag << 32) | i)
    return RunSpec(width=side, height=side, agents=int(n), params=p, discretization=D,
                   heading_mode=heading_mode, init_seed=(tag << 32) | i)


def batched(k=512, seed=3, n_min=1_000, n_max=262_144):
    """k replicas with n log-uniform in [n_min, n_max], side for constant
    density, weights U[0.5, 2], jump U[0.3, 1]; replica i keyed (seed, i)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        out.append(_run(rng, n, seed, i))
    return out


@dataclass
class Request:
    arrival: float  # seconds from the start of the stream
    spec: RunSpec
    steps: int


def serving(rate=200.0, duration=1.0, seed=4, n_min=1_000, n_max=1_000_000, s_min=50,
            s_max=2000):
    """Poisson arrivals at `rate` runs/s for `duration` s: agents log-uniform
    in [n_min, n_max], steps uniform in {s_min..s_max}, sampled weights."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    i = 0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        n = int(np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        steps = int(rng.integers(s_min, s_max + 1))
        out.append(Request(t, _run(rng, n, seed, i), steps))
        i += 1
    return out


def clustered(n=8_388_608, side=40_960.0, blobs=64, sigma=200.0, uniform_frac=0.2, seed=2):
    """Positions of the clustered-density workload: a fraction `uniform_frac`
    uniform on the torus, the rest drawn from `blobs` Gaussian blobs (centres
    uniform, standard deviation `sigma`) wrapped onto the torus with
    toroidal_transform; headings unit vectors at uniform angles.  Returns
    float32 arrays (x, y, hx, hy)."""
    rng = np.random.default_rng(seed)
    nu = int(round(n * uniform_frac))
    nb = n - nu
    centres = rng.uniform(0.0, side, size=(blobs, 2))
    which = rng.integers(0, blobs, size=nb)
    pts = centres[which] + rng.normal(0.0, sigma, size=(nb, 2))
    pts = np.concatenate([pts, rng.uniform(0.0, side, size=(nu, 2))])
    rng.shuffle(pts)
    pts = np.mod(pts, side)
    x = pts[:, 0].astype(np.float32)
    y = pts[:, 1].astype(np.float32)
    # float64 -> float32 rounding can land exactly on `side`; wrap it as
    # toroidal_transform does (field_2d.rs:1004-1014)
    x[x >= np.float32(side)] -= np.float32(side)
    y[y >= np.float32(side)] -= np.float32(side)
    a = rng.uniform(0.0, 2.0 * np.pi, size=n)
    return x, y, np.cos(a).astype(np.float32), np.sin(a).astype(np.float32)
