"""The CPU oracle, pinned without the reference (which cannot be built here).

What pins what:
  - Philox4x32-10: the published Random123 known-answer vectors
    (tests/golden/philox_kat.json).
  - Field2D: the reference's own Field2D tests (tests/engine/field_2d.rs:30-247)
    ported; an independent numpy restatement of the query rules evaluated on
    all pairs (so a wrong clamp, window or check_circle case fails); the
    textbook radius query as a superset (the reference may drop, never add);
    relax-query superset; helper identities.
  - Bird::step: closed forms of each term in isolation (weights zeroed), the
    double division of consistency, the self exclusion, |last_d| = jump.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
f32 = np.float32


# ---- Philox -----------------------------------------------------------------

def test_philox_known_answers(oracle):
    with open(os.path.join(HERE, "golden", "philox_kat.json")) as fh:
        kats = json.load(fh)["vectors"]
    assert len(kats) == 3
    for v in kats:
        got = oracle.philox4x32_10([int(c, 16) for c in v["ctr"]], [int(k, 16) for k in v["key"]])
        assert [f"{int(x):08x}" for x in got] == v["out"]


def test_u01_map(oracle):
    L = oracle.lib()
    assert L.or_u01(0) == 0.0
    assert L.or_u01(0xFFFFFFFF) == f32(1.0 - 2.0 ** -24)
    assert L.or_u01(0x80000000) == 0.5
    assert L.or_u01(0x000000FF) == 0.0   # only the top 24 bits count


def test_init_uniform_properties(oracle):
    x, y, hx, hy = oracle.init_uniform(20000, 37.0, 11.0, 3, 1)
    assert np.all((x >= 0) & (x < 37.0)) and np.all((y >= 0) & (y < 11.0))
    norm = np.hypot(hx.astype(np.float64), hy.astype(np.float64))
    assert np.allclose(norm, 1.0, atol=2e-7)
    ang = np.arctan2(hy, hx)
    hist, _ = np.histogram(ang, bins=8, range=(-np.pi, np.pi))
    assert hist.min() > 0.9 * hist.mean()  # uniform angle
    x0, y0, hx0, hy0 = oracle.init_uniform(100, 10.0, 10.0, 3, 0)
    assert np.all(hx0 == 0) and np.all(hy0 == 0)  # state.rs:47


# ---- field_2d.rs helpers -----------------------------------------------------

def test_helper_identities(oracle):
    L = oracle.lib()
    # (n % size) + size for negative n: -4 maps to 4, not 0 (field_2d.rs:926-932)
    for n, size, want in [(5, 4, 1), (-1, 4, 3), (-4, 4, 4), (0, 7, 0), (13, 7, 6)]:
        assert L.or_t_transform(n, size) == want
    w = 200.0
    assert L.or_toroidal_transform(-0.25, w) == f32(199.75)
    assert L.or_toroidal_transform(200.5, w) == f32(0.5)
    assert L.or_toroidal_transform(17.0, w) == 17.0
    rng = np.random.default_rng(1)
    for a, b in rng.uniform(-10, 210, (500, 2)).astype(np.float32):
        d1 = L.or_toroidal_distance(a, b, w)
        d2 = L.or_toroidal_distance(b, a, w)
        assert d1 == -d2
        assert abs(d1) <= 100.0
    assert L.or_f32_to_i32(float("nan")) == 0
    assert L.or_f32_to_i32(3.0e9) == 2**31 - 1
    assert L.or_f32_to_i32(-3.0e9) == -2**31


def test_check_circle_truth_table(oracle):
    L = oracle.lib()
    # cell (0,0) of side 1 in a 10x10 torus: corners (0,0),(0,1),(1,0),(1,1)
    assert L.or_check_circle(0, 0, 1.0, 10.0, 10.0, 0.5, 0.5, 0.8, 1) == 1   # all corners in
    assert L.or_check_circle(0, 0, 1.0, 10.0, 10.0, 5.0, 5.0, 1.0, 1) == -1  # all out
    assert L.or_check_circle(0, 0, 1.0, 10.0, 10.0, 0.0, 0.0, 1.0, 1) == 0   # three of four
    # a circle cutting an edge but holding no corner is dropped (field_2d.rs:963-968)
    assert L.or_check_circle(0, 0, 1.0, 10.0, 10.0, 0.5, 1.3, 0.4, 0) == -1


# ---- the query rules, restated independently in numpy ------------------------

def rule_query(xs, ys, qx, qy, dist, w, h, d, tor):
    """field_2d.rs:386-440 evaluated on all agents at once (float32)."""
    xs = xs.astype(f32)
    ys = ys.astype(f32)
    qx, qy, dist, w, h, d = (f32(v) for v in (qx, qy, dist, w, h, d))
    if not dist > 0:
        return np.zeros(0, dtype=int)
    dd = int(np.floor(dist / d))
    cx, cy = int(np.floor(qx / d)), int(np.floor(qy / d))
    max_x, max_y = int(np.ceil(w / d)), int(np.ceil(h / d))
    lo_i, hi_i, lo_j, hi_j = cx - dd, cx + dd, cy - dd, cy + dd
    if tor:
        lo_i, hi_i, lo_j, hi_j = max(0, lo_i), min(hi_i, max_x - 1), max(0, lo_j), min(hi_j, max_y - 1)

    def tdist(a, b, dim):
        a = np.asarray(a, f32)
        b = np.asarray(b, f32)
        direct = (a - b).astype(f32)
        ta = np.where((a >= 0) & (a < dim), a, np.fmod(a, dim)).astype(f32)
        ta = np.where(ta < 0, ta + dim, ta).astype(f32)
        tb = np.where((b >= 0) & (b < dim), b, np.fmod(b, dim)).astype(f32)
        tb = np.where(tb < 0, tb + dim, tb).astype(f32)
        dd_ = (ta - tb).astype(f32)
        wrapped = np.where(dd_ * f32(2) > dim, dd_ - dim, np.where(dd_ * f32(2) < -dim, dd_ + dim, dd_))
        return np.where(np.abs(direct) <= dim / f32(2), direct, wrapped).astype(f32)

    def dist_to(px, py):
        if tor:
            dx, dy = tdist(px, qx, w), tdist(py, qy, h)
        else:
            dx, dy = (np.asarray(px, f32) - qx).astype(f32), (np.asarray(py, f32) - qy).astype(f32)
        return np.sqrt((dx * dx + dy * dy).astype(f32)).astype(f32)

    ex, ey = np.floor(xs / d).astype(int), np.floor(ys / d).astype(int)
    out = []
    for i in range(lo_i, hi_i + 1):
        for j in range(lo_j, hi_j + 1):
            bx, by = i % max_x, j % max_y
            nw = (f32(bx) * d, f32(by) * d)
            ne = (nw[0], min(f32(nw[1] + d), h))
            sw = (min(f32(nw[0] + d), w), nw[1])
            se = (sw[0], ne[1])
            cd = [dist_to(*c) for c in (nw, ne, sw, se)]
            check = 1 if all(c <= dist for c in cd) else (-1 if all(c > dist for c in cd) else 0)
            members = np.nonzero((ex == bx) & (ey == by))[0]
            for m in members:  # bag order = insertion (= id) order
                if check == 1 or (check == 0 and dist_to(xs[m], ys[m]) <= dist):
                    out.append(m)
    return np.array(out, dtype=int)


@pytest.mark.parametrize("tor", [True, False])
@pytest.mark.parametrize("side,d,dist", [(50.0, 2.0, 3.7), (200.0, 10 / 1.5, 10.0), (30.0, 1.3, 2.6)])
def test_query_rules_restated(oracle, tor, side, d, dist):
    rng = np.random.default_rng(7)
    n = 600
    xs = rng.uniform(0, side, n).astype(f32)
    ys = rng.uniform(0, side, n).astype(f32)
    # cell-boundary and far-edge positions
    xs[:40] = (np.floor(rng.uniform(0, side / d, 40)) * f32(d)).astype(f32)
    ys[40:60] = np.nextafter(f32(side), f32(0))
    fld = oracle.Field2D(side, side, d, tor)
    for i in range(n):
        fld.set_object_location(i, float(xs[i]), float(ys[i]))
    fld.lazy_update()
    for q in range(80):
        qx, qy = (float(v) for v in rng.uniform(0, side, 2).astype(f32))
        got = [b[0] for b in fld.get_neighbors_within_distance(qx, qy, dist)]
        want = list(rule_query(xs, ys, qx, qy, dist, side, side, d, tor))
        assert got == want
        if tor:
            # never more than the textbook radius query
            ddx = np.minimum(np.abs(xs - qx), side - np.abs(xs - qx))
            ddy = np.minimum(np.abs(ys - qy), side - np.abs(ys - qy))
            exact = set(np.nonzero(np.hypot(ddx, ddy) <= dist * (1 + 1e-5))[0])
            assert set(got) <= exact
        relax = {b[0] for b in fld.get_neighbors_within_relax_distance(qx, qy, dist)}
        assert set(got) <= relax


# ---- tests/engine/field_2d.rs on the oracle ------------------------------------

def test_reference_field_tests_on_oracle(oracle):
    W = H = 10.0
    Dd = 0.5
    # field_2d_neighbors (:57-117)
    for fly in range(5, 10):
        f = oracle.Field2D(W, H, Dd, True)
        f.set_object_location(1, 0.0, 0.0)
        f.set_object_location(2, 0.0, 0.0)
        f.lazy_update()
        assert f.nagents == 2
        assert len(f.get_neighbors_within_distance(5.0, 5.0, 1.0)) == 0
        f.set_object_location(1, fly, fly)
        f.set_object_location(2, 0.0, 0.0)
        f.lazy_update()
        assert [b[0] for b in f.get_neighbors_within_distance(fly, fly, 1.0)] == [1]
        assert [b[0] for b in f.get_neighbors_within_distance(0.0, 0.0, 1.0)] == [2]
        assert sorted(b[0] for b in f.get_neighbors_within_distance(5.0, 5.0, 10.0)) == [1, 2]
    # field_2d_gets (:124-178)
    f = oracle.Field2D(W, H, Dd, True)
    f.set_object_location(1, 0.0, 0.0)
    f.set_object_location(2, 5.0, 5.0)
    f.set_object_location(3, 5.0, 5.0)
    f.lazy_update()
    assert sorted(b[0] for b in f.get_objects(5.0, 5.0)) == [2, 3]
    assert len(f.get_objects(10.0, 0.0)) == 0
    assert f.num_objects_at_location(5.0, 5.0) == 2
    f.set_object_location(4, 0.0, 0.0)
    assert len(f.get_objects_unbuffered(0.0, 0.0)) == 1
    f.remove_object_location(4, 0.0, 0.0)
    assert len(f.get_objects_unbuffered(0.0, 0.0)) == 0
    # field_2d_bags (:185-210)
    f = oracle.Field2D(W, H, Dd, True)
    assert f.num_empty_bags() == 21 * 21
    for i, (x, y) in enumerate([(0, 0), (0, 0), (4, 4)]):
        f.set_object_location(i + 1, x, y)
    f.lazy_update()
    assert f.num_empty_bags() == 21 * 21 - 2


def test_reference_single_step_on_oracle(oracle):
    # field_2d_single_step (:30-50): 10 birds in the 10x10 flockers field
    x, y, hx, hy = oracle.init_uniform(10, 10.0, 10.0, 4, 0)
    fl = oracle.Flocker(10.0, 10.0, 0.5, True, x, y, hx, hy, oracle.params(radius=10.0))
    fl.step(1)
    s = fl.state()
    fld = oracle.Field2D(10.0, 10.0, 0.5, True)
    for i in range(10):
        fld.set_object_location(i, float(s[0, i]), float(s[1, i]))
    fld.lazy_update()
    assert len(fld.get_neighbors_within_distance(5.0, 5.0, 10.0)) == 10


# ---- Bird::step closed forms -----------------------------------------------------

def one_step(oracle, xs, ys, hxs, hys, side=100.0, d=10 / 1.5, **w):
    kw = dict(cohesion=0.0, avoidance=0.0, consistency=0.0, randomness=0.0, momentum=0.0)
    kw.update(w)
    fl = oracle.Flocker(side, side, d, True, np.array(xs, f32), np.array(ys, f32),
                        np.array(hxs, f32), np.array(hys, f32), oracle.params(**kw))
    fl.step(1)
    return fl.state()


def test_momentum_only_keeps_direction(oracle):
    s = one_step(oracle, [50.0], [50.0], [3.0], [4.0], momentum=1.0)
    assert np.allclose(s[2:, 0], [0.6 * 0.7, 0.8 * 0.7], atol=1e-7)
    assert np.allclose(s[:2, 0], [50.0 + 0.42, 50.0 + 0.56], atol=1e-5)


def test_cohesion_pulls_together_and_avoidance_pushes_apart(oracle):
    # bird 0 at x=50, bird 1 at x=53: dx = toroidal_distance(self, other) = -3
    s = one_step(oracle, [50.0, 53.0], [50.0, 50.0], [0, 0], [0, 0], cohesion=1.0)
    assert s[2, 0] == f32(0.7) and s[2, 1] == f32(-0.7)      # towards each other
    s = one_step(oracle, [50.0, 53.0], [50.0, 50.0], [0, 0], [0, 0], avoidance=1.0)
    assert s[2, 0] == f32(-0.7) and s[2, 1] == f32(0.7)      # away from each other


def test_consistency_divides_twice(oracle):
    # bird 0 sees two neighbours with heading (1,0): consistency = (2,0)/2/2 =
    # (0.5, 0) (bird.rs:84-94); with momentum (0,1) the step direction is
    # (0.5, 1) (a single division would give (1, 1)).
    s = one_step(oracle, [50.0, 51.0, 50.0], [50.0, 50.0, 51.0], [0, 1, 1], [1, 0, 0],
                 consistency=1.0, momentum=1.0)
    h = s[2:, 0].astype(np.float64)
    assert np.allclose(h / np.linalg.norm(h), np.array([0.5, 1.0]) / np.hypot(0.5, 1.0), atol=1e-6)


def test_isolated_bird_draws_randomness_and_excludes_itself(oracle):
    # a lone bird finds only itself: count = 0, randomness still applied
    s = one_step(oracle, [50.0], [50.0], [0.0], [0.0], randomness=1.0)
    r1, r2 = oracle.step_uniforms(0, 0, 0)
    xr, yr = f32(r1) * f32(2) - f32(1), f32(r2) * f32(2) - f32(1)
    want = np.array([xr, yr], dtype=np.float64)
    assert np.allclose(s[2:, 0] / np.linalg.norm(s[2:, 0]), want / np.linalg.norm(want), atol=1e-6)
    assert abs(np.linalg.norm(s[2:, 0].astype(np.float64)) - 0.7) < 1e-6
