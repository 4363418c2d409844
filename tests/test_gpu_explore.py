"""Model exploration on the device (explore_parallel! / explore_sequential!,
model_exploration.rs:231-432): every run of a sweep is the run the oracle
computes for its configuration and Philox key, bit for bit, whether the
sweep batches the runs into one Ensemble or runs them one by one."""
import numpy as np
import pytest

from krabmaga_amd.explore import ExploreMode, explore, flockers_spec, write_csv, load_csv

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


INPUTS = {"agents": [600, 1500], "jump": [0.4, 0.9], "cohesion": [0.5, 1.7]}


def test_parallel_and_sequential_sweeps_agree():
    kw = dict(nstep=25, rep_conf=2, inputs=INPUTS, outputs=("polarisation", "mean_x"),
              mode=ExploreMode.Exaustive, seed=3, keep_state=True)
    par = explore(parallel=True, **kw)
    seq = explore(parallel=False, **kw)
    assert len(par) == len(seq) == 8 * 2
    for k, (p, s) in enumerate(zip(par, seq)):
        assert (p.conf_num, p.conf_rep) == (k // 2, k % 2)
        assert (s.conf_num, s.conf_rep) == (k // 2, k % 2 + 1)
        assert p.inputs == s.inputs
        assert np.array_equal(bits(p.state), bits(s.state)), f"run {k}"
        assert p.outputs == s.outputs
        assert p.step_per_sec > 0 and s.run_duration > 0
    # repetitions of one configuration differ (their own Philox keys)
    assert not np.array_equal(par[0].state, par[1].state)


def test_sweep_runs_match_the_oracle(oracle):
    inputs = {"agents": [400, 900], "width": [150.0, 190.0], "momentum": [0.7, 1.3],
              "randomness": [0.6, 1.4]}
    steps = 30
    rows = explore(steps, 2, inputs, outputs=("polarisation",), mode=ExploreMode.Matched,
                   seed=7, keep_state=True, hbm_budget=1 << 30)
    assert len(rows) == 4
    for k, r in enumerate(rows):
        s = flockers_spec(r.inputs, (7 << 32) | k)
        x, y, hx, hy = oracle.init_uniform(s.agents, s.width, s.height, s.init_seed, s.heading_mode)
        p = s.params
        fl = oracle.Flocker(s.width, s.height, s.discretization, True, x, y, hx, hy,
                            oracle.params(cohesion=p.cohesion, avoidance=p.avoidance,
                                          consistency=p.consistency, randomness=p.randomness,
                                          momentum=p.momentum, jump=p.jump, radius=p.radius,
                                          seed=p.seed))
        fl.step(steps)
        assert np.array_equal(bits(r.state), bits(fl.state())), f"run {k} ({r.inputs})"


def test_small_budget_splits_into_batches_with_the_same_bits(tmp_path):
    kw = dict(nstep=12, rep_conf=3, inputs={"agents": [300, 700, 1200]},
              outputs=("polarisation",), seed=1, keep_state=True)
    one = explore(**kw)
    many = explore(hbm_budget=600_000, **kw)
    for a, b in zip(one, many):
        assert np.array_equal(bits(a.state), bits(b.state))
    write_csv(str(tmp_path / "r"), many)
    (pol,) = load_csv(str(tmp_path / "r.csv"), ("x", int), ("y", int), ("agents", int),
                      ("polarisation", float))[3:]
    assert pol == [r.outputs["polarisation"] for r in many]


def test_zero_steps_reports_the_initial_flock(oracle):
    for parallel in (True, False):
        (r,) = explore(0, 1, {"agents": [500]}, seed=5, parallel=parallel, keep_state=True)
        s = flockers_spec(r.inputs, 5 << 32)
        init = np.stack(oracle.init_uniform(s.agents, s.width, s.height, s.init_seed,
                                            s.heading_mode))
        assert np.array_equal(bits(r.state), bits(init)), f"parallel={parallel}"
        assert r.step_per_sec == 0.0


def test_end_condition_stops_each_run_at_its_step(oracle):
    """State::end_condition after every schedule.step (model_exploration.rs:
    173-177): run k stops after 3 + 2k steps (one condition reads the state);
    its birds are the oracle's after that many steps and its Step per sec
    counts those steps."""
    inputs = {"agents": [300, 500, 800]}
    stop = {}

    def end_condition(step, state, spec):
        k = spec.init_seed & 0xFFFFFFFF
        if k == 1:
            assert state().shape == (4, spec.agents)
        stop[k] = step
        return step >= 3 + 2 * k

    for parallel in (True, False):
        rows = explore(20, 1, inputs, mode=ExploreMode.Matched, seed=9, keep_state=True,
                       parallel=parallel, end_condition=end_condition)
        for k, r in enumerate(rows):
            assert stop[k] == 3 + 2 * k, f"parallel={parallel} run {k}"
            assert r.step_per_sec == pytest.approx((3 + 2 * k) / r.run_duration, rel=1e-6)
            s = flockers_spec(r.inputs, (9 << 32) | k)
            x, y, hx, hy = oracle.init_uniform(s.agents, s.width, s.height, s.init_seed,
                                                s.heading_mode)
            p = s.params
            fl = oracle.Flocker(s.width, s.height, s.discretization, True, x, y, hx, hy,
                                oracle.params(cohesion=p.cohesion, avoidance=p.avoidance,
                                              consistency=p.consistency,
                                              randomness=p.randomness, momentum=p.momentum,
                                              jump=p.jump, radius=p.radius, seed=p.seed))
            fl.step(3 + 2 * k)
            assert np.array_equal(bits(r.state), bits(fl.state())), \
                f"parallel={parallel} run {k}"
