"""Flockers step on the GPU against the CPU oracle.

Bar (integer/index work bit-exact, f32 within 1e-5 of the domain width):
  - neighbour id lists equal, element for element, to
    Field2D::get_neighbors_within_distance (field_2d.rs:386-440);
  - positions and headings after one Schedule::step within 1e-5 * width --
    in fact bit-identical, because the device follows the same f32 operation
    sequence in the same neighbour order, so multi-step runs are compared
    bit for bit as well;
  - at the 1M-boid size, where the oracle would take minutes, invariants that
    do not depend on size.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = 10.0 / 1.5


def gpu_field(n, side, d, seed, heading_mode=1, bin_rows=0, toroidal=True):
    from krabmaga_amd import Field2D
    f = Field2D(side, side, d, toroidal, n, bin_rows=bin_rows)
    f.init_uniform(n, seed, heading_mode)
    return f


def oracle_flocker(O, n, side, d, seed, heading_mode=1, toroidal=True, **kw):
    x, y, hx, hy = O.init_uniform(n, side, side, seed, heading_mode)
    return O.Flocker(side, side, d, toroidal, x, y, hx, hy, O.params(seed=seed, **kw))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def test_init_matches_oracle(oracle):
    f = gpu_field(5000, 500.0, D, seed=11)
    got, ids = f.state_by_id(1)
    assert np.array_equal(ids, np.arange(5000))
    ref = np.stack(oracle.init_uniform(5000, 500.0, 500.0, 11, 1))
    assert np.array_equal(bits(got), bits(ref))


@pytest.mark.parametrize("n,side,d,radius,bin_rows", [
    (1000, 200.0, D, 10.0, 0),      # flockers config, 3x3 windows
    (1000, 200.0, D, 10.0, 4),
    (3000, 120.0, 4.0, 10.0, 0),    # 5x5 windows (disc_dist 2)
    (2000, 100.0, 0.5, 3.0, 0),     # large windows, fine cells
    (500, 37.5, 2.5, 5.0, 0),       # grid with a partial padding cell
])
def test_neighbor_sets_bit_exact(oracle, n, side, d, radius, bin_rows):
    from krabmaga_amd import Field2D
    x, y, hx, hy = oracle.init_uniform(n, side, side, 5, 1)
    f = Field2D(side, side, d, True, n, bin_rows=bin_rows)
    f.set_object_locations(np.arange(n), x, y, hx, hy)
    f.lazy_update()
    of = oracle.Field2D(side, side, d, True)
    for i in range(n):
        of.set_object_location(i, float(x[i]), float(y[i]))
    of.lazy_update()
    storage, row_ptr, ids = f.neighbor_sets(radius)
    a = f.read_agents(0)
    for k in range(len(storage)):
        want = [b[0] for b in of.get_neighbors_within_distance(float(a["x"][k]), float(a["y"][k]), radius)]
        got = list(ids[row_ptr[k]:row_ptr[k + 1]])
        assert got == want, (k, storage[k])


def test_single_queries_bit_exact_including_non_toroidal(oracle):
    from krabmaga_amd import Field2D
    rng = np.random.default_rng(3)
    for tor in (True, False):
        n, side, d = 400, 50.0, 2.0
        x = rng.uniform(0, side, n).astype(np.float32)
        y = rng.uniform(0, side, n).astype(np.float32)
        f = Field2D(side, side, d, tor, n)
        f.set_object_locations(np.arange(n), x, y)
        f.lazy_update()
        of = oracle.Field2D(side, side, d, tor)
        for i in range(n):
            of.set_object_location(i, float(x[i]), float(y[i]))
        of.lazy_update()
        for q in range(60):
            qx, qy = (float(v) for v in rng.uniform(0, side, 2).astype(np.float32))
            r = float(rng.choice([0.5, 2.0, 3.7, 6.0, 13.0]))
            assert list(f.get_neighbors_within_distance(qx, qy, r)) == \
                [b[0] for b in of.get_neighbors_within_distance(qx, qy, r)]
            assert list(f.get_neighbors_within_relax_distance(qx, qy, r)) == \
                [b[0] for b in of.get_neighbors_within_relax_distance(qx, qy, r)]


@pytest.mark.parametrize("heading_mode", [0, 1])
def test_single_step_within_tolerance(oracle, heading_mode):
    from krabmaga_amd import FlockParams
    n, side = 1000, 200.0
    f = gpu_field(n, side, D, seed=0, heading_mode=heading_mode)
    f.schedule_run(FlockParams(seed=0), 0, 1, use_graph=False)
    got, _ = f.state_by_id(0)
    fl = oracle_flocker(oracle, n, side, D, 0, heading_mode)
    fl.step(1)
    ref = fl.state()
    tol = 1e-5 * side
    assert np.max(np.abs(got - ref)) <= tol
    assert np.array_equal(bits(got), bits(ref))


@pytest.mark.parametrize("n,side,steps,seed", [
    (1000, 200.0, 200, 0),      # the flockers CPU configuration, whole run
    (20000, 900.0, 20, 1),
    (50000, 4472.0, 10, 2),     # sparse: 0.0025 birds per unit^2, the 1M-boid field's density
])
def test_multi_step_bit_exact(oracle, n, side, steps, seed):
    from krabmaga_amd import FlockParams
    f = gpu_field(n, side, D, seed=seed)
    f.schedule_run(FlockParams(seed=seed), 0, steps, use_graph=True)
    got, ids = f.state_by_id(0)
    assert np.array_equal(ids, np.arange(n))
    fl = oracle_flocker(oracle, n, side, D, seed)
    fl.step(steps)
    assert np.array_equal(bits(got), bits(fl.state()))


@pytest.mark.parametrize("n,side,seed", [(1000, 200.0, 0), (60000, 4000.0, 3), (20000, 300.0, 5)])
def test_tile_resident_and_binned_paths_agree(oracle, n, side, seed):
    """The one-launch tile-resident step equals the three-kernel path and the oracle."""
    from krabmaga_amd import FlockParams
    p = FlockParams(seed=seed)
    a = gpu_field(n, side, D, seed=seed)
    b = gpu_field(n, side, D, seed=seed)
    a.schedule_run(p, 0, 12)
    b.schedule_run(p, 0, 12, binned=True)
    sa, ia = a.state_by_id(0)
    sb, ib = b.state_by_id(0)
    assert np.array_equal(ia, np.arange(n)) and np.array_equal(ib, np.arange(n))
    assert np.array_equal(bits(sa), bits(sb))
    if n <= 20000:
        fl = oracle_flocker(oracle, n, side, D, seed)
        fl.step(12)
        assert np.array_equal(bits(sa), bits(fl.state()))


def test_run_split_across_calls_and_layouts(oracle):
    """Runs resumed across calls, queries in between (layout round trips)."""
    from krabmaga_amd import FlockParams
    n, side, seed = 3000, 300.0, 8
    p = FlockParams(seed=seed)
    f = gpu_field(n, side, D, seed=seed)
    f.schedule_run(p, 0, 5)
    assert len(f.get_neighbors_within_distance(150.0, 150.0, 10.0)) >= 0  # back to bins
    f.schedule_run(p, 5, 9, use_graph=False)
    f.schedule_run(p, 14, 3, binned=True)
    f.schedule_run(p, 17, 20)
    s, _ = f.state_by_id(0)
    fl = oracle_flocker(oracle, n, side, D, seed)
    fl.step(37)
    assert np.array_equal(bits(s), bits(fl.state()))


def test_graph_and_direct_paths_agree():
    from krabmaga_amd import FlockParams
    p = FlockParams(seed=9, cohesion=1.3, jump=0.5)
    a = gpu_field(30000, 1100.0, D, seed=9)
    b = gpu_field(30000, 1100.0, D, seed=9)
    a.schedule_run(p, 0, 7, use_graph=True)
    a.schedule_run(p, 7, 6, use_graph=True)
    b.schedule_run(p, 0, 13, use_graph=False)
    sa, _ = a.state_by_id(0)
    sb, _ = b.state_by_id(0)
    assert np.array_equal(bits(sa), bits(sb))


def test_flock_step_then_lazy_update_equals_schedule_run():
    from krabmaga_amd import FlockParams
    p = FlockParams(seed=4)
    a = gpu_field(4000, 400.0, D, seed=4)
    b = gpu_field(4000, 400.0, D, seed=4)
    a.lazy_update()  # schedule.rs:349-351 at step 0
    for s in range(5):
        a.flock_step(p, s)
        a.lazy_update()
    b.schedule_run(p, 0, 5, use_graph=False)
    sa, _ = a.state_by_id(0)
    sb, _ = b.state_by_id(0)
    assert np.array_equal(bits(sa), bits(sb))


@pytest.mark.parametrize("n,side,jump,seed", [
    (3000, 1200.0, 150.0, 20),  # jumps past the neighbouring tiles: counts come up short
    (6000, 200.0, 60.0, 18),    # most birds leave their tile: outboxes overflow
    (20000, 900.0, 0.7, 1),     # the gather rebins every step
])
def test_gather_rebinning_and_its_fallback_bit_exact(oracle, n, side, jump, seed):
    """Flock steps followed by lazy_update rebin by gather; the steps it cannot
    do go to the scatter and the bin sort.  Either way the run equals the oracle."""
    from krabmaga_amd import FlockParams
    p = FlockParams(seed=seed, jump=jump)
    f = gpu_field(n, side, D, seed=seed)
    f.lazy_update()  # schedule.rs:349-351 at step 0
    for s in range(6):
        f.flock_step(p, s)
        f.lazy_update()
    got, ids = f.state_by_id(0)
    assert np.array_equal(ids, np.arange(n))
    fl = oracle_flocker(oracle, n, side, D, seed, jump=jump)
    fl.step(6)
    assert np.array_equal(bits(got), bits(fl.state()))


def test_inline_global_path_equals_the_global_launch():
    """A uniform field's first step shows no tile on the global path, so later
    steps of schedule_run step such tiles inside k_flock; a dense cluster
    added afterwards sends tiles there.  kg_flock_step keeps launching
    k_flock_global: both runs agree bit for bit."""
    from krabmaga_amd import Field2D, FlockParams
    n, m, side, seed = 20000, 6000, 900.0, 12
    p = FlockParams(seed=seed)
    rng = np.random.default_rng(seed)
    cx = (400.0 + 20.0 * rng.random(m)).astype(np.float32)
    cy = (300.0 + 20.0 * rng.random(m)).astype(np.float32)
    ang = rng.random(m) * 2 * np.pi
    hx, hy = np.cos(ang).astype(np.float32), np.sin(ang).astype(np.float32)
    ids = np.arange(n, n + m, dtype=np.uint32)
    runs = []
    for by_step in (False, True):
        f = Field2D(side, side, D, True, n + m)
        f.init_uniform(n, seed, 1)
        if by_step:
            f.lazy_update()
            for s in range(4):
                f.flock_step(p, s)
                f.lazy_update()
        else:
            f.schedule_run(p, 0, 4, binned=True)
        # the write buffer holds what is set for the next step: every bird
        # so far, then the cluster
        st0, ids0 = f.state_by_id(0)
        f.set_object_locations(np.concatenate([ids0, ids]), np.concatenate([st0[0], cx]),
                               np.concatenate([st0[1], cy]), np.concatenate([st0[2], hx]),
                               np.concatenate([st0[3], hy]))
        f.lazy_update()
        if by_step:
            for s in range(4, 10):
                f.flock_step(p, s)
                f.lazy_update()
        else:
            f.schedule_run(p, 4, 6, binned=True)
        st, got_ids = f.state_by_id(0)
        assert np.array_equal(got_ids, np.arange(n + m))
        runs.append(st)
        f.close()
    assert np.array_equal(bits(runs[0]), bits(runs[1]))


@pytest.mark.parametrize("side", [20480.0, 6476.0])  # the 1M configuration; at the 1,000-boid density
def test_one_million_boids_invariants(side):
    """1,048,576 boids on a torus: properties independent of size, and the
    tile-resident and three-kernel paths agree bit for bit."""
    from krabmaga_amd import FlockParams
    n = 1 << 20
    p = FlockParams(seed=1)
    f = gpu_field(n, side, D, seed=1)
    f.schedule_run(p, 0, 20)
    s1, ids = f.state_by_id(0)
    assert np.array_equal(ids, np.arange(n))                      # N conserved, ids intact
    assert np.all((s1[0] >= 0) & (s1[0] < side))                  # toroidal_transform range
    assert np.all((s1[1] >= 0) & (s1[1] < side))
    norm = np.hypot(s1[2].astype(np.float64), s1[3].astype(np.float64))
    assert np.allclose(norm, 0.7, atol=1e-5)                      # |last_d| = jump
    g = gpu_field(n, side, D, seed=1)                             # determinism
    g.schedule_run(p, 0, 20)
    s2, _ = g.state_by_id(0)
    assert np.array_equal(bits(s1), bits(s2))
    g.close()
    h = gpu_field(n, side, D, seed=1)                             # the other device path
    h.schedule_run(p, 0, 20, binned=True)
    s3, _ = h.state_by_id(0)
    h.close()
    f.close()
    assert np.array_equal(bits(s1), bits(s3))
