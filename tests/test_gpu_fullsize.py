"""Parity at the full sizes of the headline workloads, against the oracle.

  - 1,048,576 boids on a 20,480 x 20,480 torus (0.0025 birds per unit^2),
    and the same boids on 6,476 x 6,476 (the 1,000-boid configuration's
    0.025 per unit^2): every bird's neighbour vector after binning
    (Field2D::get_neighbors_within_distance, field_2d.rs:386-440) equals the
    oracle's element for element, and two Schedule::step calls equal the
    oracle's two steps bit for bit, on the default (three-kernel) path and
    on the tile-resident step.
  - 8,388,608 boids on a 40,960 x 40,960 torus, 80 % in 64 Gaussian blobs
    (sigma 200), 20 % uniform: one Schedule::step equals the oracle's bit
    for bit, and neighbour vectors at 2,000 sampled birds equal the
    oracle's.  The oracle takes about a minute for this step.
"""
import time

import numpy as np
import pytest

from krabmaga_amd import Field2D, FlockParams, workloads

pytestmark = pytest.mark.gpu

D = 10.0 / 1.5
N1 = 1 << 20


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def say(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


@pytest.mark.parametrize("side", [20480.0, 6476.0])
def test_one_million_boids_match_oracle(oracle, side):
    O = oracle
    seed = 1
    x, y, hx, hy = O.init_uniform(N1, side, side, seed, 1)
    say(f"1M boids, side {side}: neighbour vectors")
    f = Field2D(side, side, D, True, N1)
    f.init_uniform(N1, seed, 1)
    f.lazy_update()
    storage, row_ptr, ids = f.neighbor_sets(10.0)
    a = f.read_agents(0)
    f.close()
    # the device's Flocker::init is the oracle's, bit for bit
    assert np.array_equal(bits(a["x"]), bits(x[a["id"]]))
    assert np.array_equal(bits(a["hy"]), bits(hy[a["id"]]))
    of = O.Field2D(side, side, D, True)
    of.set_object_locations(np.arange(N1, dtype=np.uint32), x, y, hx, hy)
    of.lazy_update()
    orp, oids = of.neighbor_sets(a["x"], a["y"], 10.0)
    del of
    assert np.array_equal(row_ptr.astype(np.uint64), orp)
    assert np.array_equal(ids, oids)
    say(f"1M boids, side {side}: {len(ids)} neighbour ids equal; two steps on each path")
    fl = O.Flocker(side, side, D, True, x, y, hx, hy, O.params(seed=seed))
    fl.step(2)
    ref = fl.state()
    del fl
    p = FlockParams(seed=seed)
    for tile_resident in (False, True):
        g = Field2D(side, side, D, True, N1)
        g.init_uniform(N1, seed, 1)
        g.schedule_run(p, 0, 2, tile_resident=tile_resident)
        # the default at this size is the three-kernel path
        assert g.stats()["rt_steps"] == (2 if tile_resident else 0)
        st, sid = g.state_by_id(0)
        g.close()
        assert np.array_equal(sid, np.arange(N1))
        bad = np.flatnonzero(np.any(bits(st) != bits(ref), axis=0))
        assert bad.size == 0, f"tile_resident={tile_resident}: {bad.size} birds differ, " \
                              f"first {bad[:5]}"


def test_clustered_8m_boids_one_step_matches_oracle(oracle):
    O = oracle
    n, side, seed = 8_388_608, 40_960.0, 2
    say("8.4M clustered boids: generating")
    x, y, hx, hy = workloads.clustered(n, side, blobs=64, sigma=200.0, uniform_frac=0.2,
                                       seed=seed)
    ids = np.arange(n, dtype=np.uint32)
    q = Field2D(side, side, D, True, n)
    q.set_object_locations(ids, x, y, hx, hy)
    q.lazy_update()
    of = O.Field2D(side, side, D, True)
    of.set_object_locations(ids, x, y, hx, hy)
    of.lazy_update()
    # 1,000 birds at random and 1,000 from cells holding 10 or more birds
    rng = np.random.default_rng(7)
    cell = (np.floor(x / np.float32(D)).astype(np.int64) * 6145 +
            np.floor(y / np.float32(D)).astype(np.int64))
    crowded = np.flatnonzero(np.bincount(cell)[cell] >= 10)
    sample = np.concatenate([rng.choice(n, 1000, replace=False),
                             rng.choice(crowded, 1000, replace=False)])
    orp, oids = of.neighbor_sets(x[sample], y[sample], 10.0)
    del of
    longest = 0
    for k, i in enumerate(sample):
        got = q.get_neighbors_within_distance(float(x[i]), float(y[i]), 10.0)
        want = oids[orp[k]:orp[k + 1]]
        longest = max(longest, len(want))
        assert np.array_equal(got, want), (int(i), len(got), len(want))
    q.close()
    assert longest >= 50  # the sample reaches blob cores
    say(f"8.4M clustered boids: 2,000 sampled neighbour vectors equal (longest {longest}); "
        "one step on the device")
    p = FlockParams(seed=seed)
    f = Field2D(side, side, D, True, n)
    f.set_object_locations(ids, x, y, hx, hy)
    f.schedule_run(p, 0, 1)
    st, sid = f.state_by_id(0)
    f.close()
    say("8.4M clustered boids: one step on the oracle")
    fl = O.Flocker(side, side, D, True, x, y, hx, hy, O.params(seed=seed))
    fl.step(1)
    ref = fl.state()
    del fl
    assert np.array_equal(sid, ids)
    bad = np.flatnonzero(np.any(bits(st) != bits(ref), axis=0))
    assert bad.size == 0, f"{bad.size} birds differ, first {bad[:5]}"
    say("8.4M clustered boids: one step equal bit for bit")
