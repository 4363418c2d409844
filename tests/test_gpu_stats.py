"""Flock statistics over 500 steps of the 1,000-boid flockers run.

The device reproduces the oracle's canonical order (agents stepped by id)
bit for bit for all 500 steps.  Against another agent order -- the
reference's own order is its scheduler's pop order, which nothing pins --
the trajectories decorrelate (the model is chaotic), so the comparison is
statistical over 8 seeds and steps 400-500: mean nearest-neighbour distance
within 1 %, polarisation within 3 standard errors of the seed-to-seed
spread (its spread between seeds is far wider than 1 % at 1,000 boids)."""
import numpy as np
import pytest

from krabmaga_amd import Ensemble, FlockParams, RunSpec
from krabmaga_amd.stats import mean_nn_distance, polarisation

pytestmark = pytest.mark.gpu

N, SIDE, D, STEPS, SEEDS = 1000, 200.0, 10.0 / 1.5, 500, 8


def sample_steps():
    return list(range(409, STEPS, 10))


def stats_of(state):
    x, y, hx, hy = state
    return polarisation(hx, hy), mean_nn_distance(x, y, SIDE, SIDE)


def test_flock_statistics_500_steps(oracle):
    O = oracle
    samples = sample_steps()
    gpu = {s: [] for s in range(SEEDS)}
    with Ensemble() as ens:
        ids = ens.admit([RunSpec(width=SIDE, height=SIDE, agents=N, params=FlockParams(seed=s),
                                 discretization=D, heading_mode=1, init_seed=s)
                         for s in range(SEEDS)])
        done = 0
        for t in samples:
            ens.run(t + 1 - done)
            done = t + 1
            for s, rid in enumerate(ids):
                st, _ = ens.state_by_id(rid)
                gpu[s].append(st)
    canon, shuf = [], []
    for s in range(SEEDS):
        x, y, hx, hy = O.init_uniform(N, SIDE, SIDE, s, 1)
        a = O.Flocker(SIDE, SIDE, D, True, x, y, hx, hy, O.params(seed=s))
        b = O.Flocker(SIDE, SIDE, D, True, x, y, hx, hy, O.params(seed=s))
        b.set_order(np.random.default_rng(100 + s).permutation(N).astype(np.uint32))
        done = 0
        ca, cb = [], []
        for k, t in enumerate(samples):
            a.step(t + 1 - done)
            b.step(t + 1 - done)
            done = t + 1
            sa = a.state()
            # the device follows the canonical order bit for bit
            assert np.array_equal(sa.view(np.uint32), gpu[s][k].view(np.uint32)), (s, t)
            ca.append(stats_of(sa))
            cb.append(stats_of(b.state()))
        canon.append(np.mean(ca, axis=0))
        shuf.append(np.mean(cb, axis=0))
    canon, shuf = np.array(canon), np.array(shuf)
    nn_a, nn_b = canon[:, 1].mean(), shuf[:, 1].mean()
    assert abs(nn_a - nn_b) / nn_b < 0.01, (nn_a, nn_b)
    pa, pb = canon[:, 0].mean(), shuf[:, 0].mean()
    se = np.sqrt(canon[:, 0].var(ddof=1) / SEEDS + shuf[:, 0].var(ddof=1) / SEEDS)
    assert abs(pa - pb) < 3 * se, (pa, pb, se)
    print(f"polarisation {pa:.4f} vs {pb:.4f} (se {se:.4f}); mean NN {nn_a:.4f} vs {nn_b:.4f}")
