"""The serving loop on the device: runs that arrive over time, overlap on
the GPU and leave at their own step counts produce exactly the bits they
produce alone, and the oracle's bits."""
import numpy as np
import pytest

from krabmaga_amd import Field2D
from krabmaga_amd.serve import Server, summary
from krabmaga_amd.workloads import serving

pytestmark = pytest.mark.gpu


def alone(spec, steps):
    f = Field2D(spec.width, spec.height, spec.discretization, True, spec.agents)
    f.init_uniform(spec.agents, spec.init_seed, spec.heading_mode)
    f.schedule_run(spec.params, 0, steps)
    st, ids = f.state_by_id(0)
    f.close()
    return st, ids


@pytest.mark.parametrize("budget_mb", [4096, 40])
def test_served_runs_equal_runs_alone(budget_mb):
    reqs = serving(rate=400.0, duration=0.05, seed=9, n_min=300, n_max=15000, s_min=5, s_max=70)
    assert len(reqs) >= 8
    srv = Server(budget_mb * 2**20, keep_state=True, max_tick=32)
    done = srv.serve(reqs)
    assert sorted(c.index for c in done) == list(range(len(reqs)))
    s = summary(done, srv.wall)
    assert s["p99_latency_s"] >= s["p50_latency_s"] > 0
    by_index = {c.index: c for c in done}
    for i in range(0, len(reqs), max(1, len(reqs) // 5)):
        r = reqs[i]
        got = by_index[i].state
        order = np.argsort(got["id"], kind="stable")
        st = np.stack([got["x"][order], got["y"][order], got["hx"][order], got["hy"][order]])
        ref, rids = alone(r.spec, r.steps)
        assert np.array_equal(got["id"][order], rids)
        bad = np.flatnonzero(np.any(st.view(np.uint32) != ref.view(np.uint32), axis=0))
        assert bad.size == 0, f"request {i} (n={r.spec.agents}, steps={r.steps}): {bad.size} differ"


def test_served_runs_equal_the_oracle(oracle):
    """Served runs against the oracle's Flocker (Field2D + Bird::step on the
    CPU), every run of a stream whose runs overlap on the device."""
    reqs = serving(rate=400.0, duration=0.04, seed=13, n_min=300, n_max=6000, s_min=5, s_max=60)
    assert len(reqs) >= 6
    srv = Server(4096 * 2**20, keep_state=True, max_tick=16)
    done = srv.serve(reqs)
    assert len(done) == len(reqs)
    for c in done:
        s, p = reqs[c.index].spec, reqs[c.index].spec.params
        x, y, hx, hy = oracle.init_uniform(s.agents, s.width, s.height, s.init_seed,
                                           s.heading_mode)
        fl = oracle.Flocker(s.width, s.height, s.discretization, True, x, y, hx, hy,
                            oracle.params(cohesion=p.cohesion, avoidance=p.avoidance,
                                          consistency=p.consistency, randomness=p.randomness,
                                          momentum=p.momentum, jump=p.jump, radius=p.radius,
                                          seed=p.seed))
        fl.step(c.steps)
        ref = fl.state()
        got = c.state
        order = np.argsort(got["id"], kind="stable")
        assert np.array_equal(got["id"][order], np.arange(s.agents, dtype=np.uint32))
        st = np.stack([got["x"][order], got["y"][order], got["hx"][order], got["hy"][order]])
        bad = np.flatnonzero(np.any(st.view(np.uint32) != ref.view(np.uint32), axis=0))
        assert bad.size == 0, f"request {c.index} (n={s.agents}, steps={c.steps}): {bad.size} differ"
