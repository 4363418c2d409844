"""Ensembles of runs (explore_parallel!, model_exploration.rs:354-432) on the
device: every run must produce exactly the bits it produces on its own field
(kg_schedule_run), whatever shares the ensemble with it, including runs that
join and leave on the way."""
import numpy as np
import pytest

from krabmaga_amd import Ensemble, Field2D, FlockParams, RunSpec, Tuning

pytestmark = pytest.mark.gpu

D = 10.0 / 1.5


def spec(n, seed, density=0.025, cohesion=1.0, avoidance=1.0, consistency=1.0, randomness=1.0,
         momentum=1.0, jump=0.7, heading_mode=1):
    side = float(np.sqrt(n / density))
    p = FlockParams(cohesion=cohesion, avoidance=avoidance, consistency=consistency,
                    randomness=randomness, momentum=momentum, jump=jump, radius=10.0,
                    seed=seed * 7919 + 3)
    return RunSpec(width=side, height=side, agents=n, params=p, heading_mode=heading_mode,
                   init_seed=seed)


def alone(s, steps):
    f = Field2D(s.width, s.height, s.discretization, True, s.agents)
    f.init_uniform(s.agents, s.init_seed, s.heading_mode)
    f.schedule_run(s.params, 0, steps)
    st, ids = f.state_by_id(0)
    f.close()
    return st, ids


def assert_same(ens, rid, s, steps, label):
    st, ids = ens.state_by_id(rid)
    ref, rids = alone(s, steps)
    assert np.array_equal(ids, np.arange(s.agents, dtype=np.uint32)), label
    assert np.array_equal(rids, ids), label
    bad = np.flatnonzero(np.any(st.view(np.uint32) != ref.view(np.uint32), axis=0))
    assert bad.size == 0, f"{label}: {bad.size} birds differ from the run alone, first {bad[:5]}"


def uneven_specs(k, base_seed=11):
    rng = np.random.default_rng(base_seed)
    out = []
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(300), np.log(30000))))
        w = rng.uniform(0.5, 2.0, size=5)
        out.append(spec(n, base_seed + i, cohesion=w[0], avoidance=w[1], consistency=w[2],
                        randomness=w[3], momentum=w[4], jump=float(rng.uniform(0.3, 1.0)),
                        heading_mode=int(i % 2)))
    return out


PATHS = [{"tile_resident": True}, {"binned": True}]


@pytest.mark.parametrize("path", PATHS, ids=["tile_resident", "binned"])
def test_ensemble_matches_runs_alone(path):
    specs = uneven_specs(6)
    with Ensemble() as ens:
        ids = ens.admit(specs)
        ens.run(40, **path)
        ens.sync()
        st = ens.stats()
        assert st["runs"] == 6 and st["agents"] == sum(s.agents for s in specs)
        for rid, s in zip(ids, specs):
            assert ens.steps(rid) == 40
            assert_same(ens, rid, s, 40, f"run {rid} (n={s.agents})")


@pytest.mark.parametrize("path", PATHS, ids=["tile_resident", "binned"])
def test_ensemble_direct_launches_match_graphs(path):
    specs = uneven_specs(3, base_seed=40)
    with Ensemble() as a, Ensemble() as b:
        ia = a.admit(specs)
        ib = b.admit(specs)
        a.run(27, use_graph=True, **path)
        b.run(27, use_graph=False, **path)
        for x, y in zip(ia, ib):
            sa, _ = a.state_by_id(x)
            sb, _ = b.state_by_id(y)
            assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))


@pytest.mark.parametrize("path", PATHS, ids=["tile_resident", "binned"])
def test_ensemble_runs_join_and_leave(path):
    s_a, s_b, s_c, s_d = uneven_specs(4, base_seed=70)
    with Ensemble() as ens:
        a, b = ens.admit([s_a, s_b])
        ens.run(20, **path)
        (c,) = ens.admit([s_c])
        ens.run(15, **path)
        assert ens.steps(a) == 35 and ens.steps(c) == 15
        assert_same(ens, a, s_a, 35, "a before leaving")
        ens.retire(a)
        (d,) = ens.admit([s_d])
        ens.run(10, **path)
        assert_same(ens, b, s_b, 45, "b")
        assert_same(ens, c, s_c, 25, "c")
        assert_same(ens, d, s_d, 10, "d")
        assert ens.stats()["builds"] >= 3
        # retiring the smallest run (at most a third of the tiles) keeps it
        # as a dead replica: no rebuild
        live = sorted([(s_b.agents, b, s_b, 45), (s_c.agents, c, s_c, 25), (s_d.agents, d, s_d, 10)])
        builds = ens.stats()["builds"]
        ens.retire(live[0][1])
        ens.run(9, **path)
        assert ens.stats()["builds"] == builds and ens.stats()["runs"] == 2
        for _, rid, s, k in live[1:]:
            assert_same(ens, rid, s, k + 9, f"run {rid} after the smallest left")
        # retiring the largest as well leaves dead tiles past a third of
        # the layout: the next run compacts it
        tiles = ens.stats()["tiles"]
        ens.retire(live[2][1])
        ens.run(5, **path)
        st = ens.stats()
        assert st["builds"] == builds + 1 and st["runs"] == 1 and st["tiles"] < tiles
        _, rid, s, k = live[1]
        assert_same(ens, rid, s, k + 14, "last run after compaction")


@pytest.mark.parametrize("path", PATHS, ids=["tile_resident", "binned"])
def test_ensemble_read_after_retire_and_admit(path):
    # a read that rebuilds the layout (an admission is pending) while a
    # retired run still sits in it: the read must address the run by id
    # after the rebuild has dropped the retired one
    s_a, s_b, s_c, s_d = uneven_specs(4, base_seed=130)
    with Ensemble() as ens:
        a, b, c = ens.admit([s_a, s_b, s_c])
        ens.run(10, **path)
        ens.retire(a)
        (d,) = ens.admit([s_d])
        assert_same(ens, c, s_c, 10, "c read during the rebuild")
        ens.run(5, **path)
        assert_same(ens, b, s_b, 15, "b")
        assert_same(ens, d, s_d, 5, "d")
        with pytest.raises(Exception):
            ens.read(a)


def test_ensemble_tuning_does_not_change_results():
    specs = uneven_specs(3, base_seed=90)
    with Ensemble(Tuning(tile_cols=8, tile_rows=32, bin_rows=8)) as ens:
        ids = ens.admit(specs)
        ens.run(25)
        for rid, s in zip(ids, specs):
            assert_same(ens, rid, s, 25, f"tuned run {rid}")


def test_ensemble_shortens_tiles_that_fit_no_lds_class():
    # 20,000 birds on a 200 x 200 torus (22 per cell) with 4 x 64-cell tiles:
    # about 5,600 birds a tile, beyond every LDS class, so the build halves
    # the tile height until the plan fits; the run keeps its bits
    s = spec(20000, 150, density=0.5)
    with Ensemble(Tuning(tile_cols=4, tile_rows=64, bin_rows=8)) as ens:
        (rid,) = ens.admit([s])
        ens.run(12)
        assert ens.stats()["tiles"] > 8  # 4 x 64 tiles would be 8
        assert_same(ens, rid, s, 12, "dense run with shortened tiles")


def test_ensemble_binned_global_tiles_among_many_birds():
    # 560,000 uniform birds make more runs of 256 than the bin-grouped step
    # gives k_flock_global work-groups, so they loop over the runs; the dense
    # run's tiles (about 22 birds a cell) exceed every LDS capacity and take
    # the global path, and overflow the gather, in every step
    specs = [spec(560000, 171), spec(20000, 172, density=0.5)]
    with Ensemble() as ens:
        ids = ens.admit(specs)
        ens.run(8, binned=True)
        for rid, s in zip(ids, specs):
            assert ens.steps(rid) == 8
            assert_same(ens, rid, s, 8, f"run {rid} (n={s.agents})")


def test_ensemble_rejects_non_toroidal_runs():
    s = spec(500, 1)
    s_bad = RunSpec(width=s.width, height=s.height, agents=500, params=s.params)
    c = s_bad.c()
    c.toroidal = 0
    from krabmaga_amd import _abi
    import ctypes
    with Ensemble() as ens:
        rc = _abi.lib().kg_ensemble_admit(ens._h, ctypes.byref(c), 1, None)
        assert rc == _abi.KG_E_INVALID


def test_ensemble_switches_between_layouts():
    # steps on one form, then the other, then back: every run keeps its bits
    specs = uneven_specs(4, base_seed=160)
    with Ensemble() as ens:
        ids = ens.admit(specs)
        ens.run(7, tile_resident=True)
        ens.run(6, binned=True)
        ens.run(5, tile_resident=True)
        ens.run(3, binned=True)
        assert ens.stats()["builds"] >= 4
        for rid, s in zip(ids, specs):
            assert ens.steps(rid) == 21
            assert_same(ens, rid, s, 21, f"run {rid} (n={s.agents})")
