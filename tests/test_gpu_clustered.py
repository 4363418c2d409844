"""Clustered density: Gaussian blobs as dense as the 8,388,608-boid
clustered workload's (peak about 16 birds per cell) on a smaller torus, so
the oracle finishes in seconds.  Tiles this full overflow every LDS working
set the device plans, which exercises the fallbacks; the results must still
be the oracle's bits."""
import numpy as np
import pytest

from krabmaga_amd import Field2D, FlockParams, workloads

pytestmark = pytest.mark.gpu

D = 10.0 / 1.5


@pytest.mark.parametrize("binned", [False, True])
def test_dense_blobs_match_oracle(oracle, binned):
    O = oracle
    n, side, steps = 131072, 10240.0, 2
    x, y, hx, hy = workloads.clustered(n, side, blobs=2, sigma=150.0, uniform_frac=0.2, seed=5)
    cx = np.floor(x / np.float32(D)).astype(np.int64)
    cy = np.floor(y / np.float32(D)).astype(np.int64)
    assert np.bincount(cx * 4096 + cy).max() >= 20  # blob cores are dense
    p = FlockParams(seed=5)
    f = Field2D(side, side, D, True, n)
    f.set_object_locations(np.arange(n, dtype=np.uint32), x, y, hx, hy)
    f.schedule_run(p, 0, steps, binned=binned)
    st, ids = f.state_by_id(0)
    f.close()
    fl = O.Flocker(side, side, D, True, x, y, hx, hy, O.params(seed=5))
    fl.step(steps)
    ref = fl.state()
    assert np.array_equal(ids, np.arange(n, dtype=np.uint32))
    bad = np.flatnonzero(np.any(st.view(np.uint32) != ref.view(np.uint32), axis=0))
    assert bad.size == 0, f"{bad.size} birds differ, first {bad[:5]}"


@pytest.mark.parametrize("binned", [False, True])
def test_crowded_cells_match_oracle(oracle, binned):
    """A few single cells holding hundreds to thousands of birds (bins far
    wider than a wave, and wider than one LDS pass of the bin sort) in a
    sparse field: two steps and the neighbour lists equal the oracle's."""
    O = oracle
    rng = np.random.default_rng(21)
    side = 400.0
    crowds = [(100.0, 100.0, 3000), (250.0, 300.0, 700), (33.0, 370.0, 90)]
    xs, ys = [rng.uniform(0.0, side, 1500)], [rng.uniform(0.0, side, 1500)]
    for cx, cy, m in crowds:
        # inside one cell of side D around (cx, cy)
        x0 = np.floor(cx / D) * D
        y0 = np.floor(cy / D) * D
        xs.append(x0 + rng.uniform(0.05, 0.95, m) * D)
        ys.append(y0 + rng.uniform(0.05, 0.95, m) * D)
    x = np.concatenate(xs).astype(np.float32)
    y = np.concatenate(ys).astype(np.float32)
    n = x.size
    a = rng.uniform(0.0, 2.0 * np.pi, n)
    hx = np.cos(a).astype(np.float32)
    hy = np.sin(a).astype(np.float32)
    perm = rng.permutation(n).astype(np.uint32)  # ids unrelated to position
    # neighbour lists straight after binning
    q = Field2D(side, side, D, True, n)
    q.set_object_locations(perm, x, y, hx, hy)
    q.lazy_update()
    # the device's bags are in ascending id, the oracle's in insertion order
    of = O.Field2D(side, side, D, True)
    for i in np.argsort(perm):
        of.set_object_location(int(perm[i]), float(x[i]), float(y[i]))
    of.lazy_update()
    for qx, qy in [(100.0, 100.0), (250.0, 300.0), (33.0, 370.0), (5.0, 5.0)]:
        want = [b[0] for b in of.get_neighbors_within_distance(qx, qy, 10.0)]
        assert list(q.get_neighbors_within_distance(qx, qy, 10.0)) == want
    q.close()
    p = FlockParams(seed=21)
    f = Field2D(side, side, D, True, n)
    f.set_object_locations(perm, x, y, hx, hy)
    f.schedule_run(p, 0, 2, binned=binned)
    st, ids = f.state_by_id(0)
    f.close()
    order = np.argsort(perm)
    fl = O.Flocker(side, side, D, True, x[order], y[order], hx[order], hy[order],
                   O.params(seed=21))
    fl.step(2)
    ref = fl.state()
    assert np.array_equal(ids, np.arange(n, dtype=np.uint32))
    bad = np.flatnonzero(np.any(st.view(np.uint32) != ref.view(np.uint32), axis=0))
    assert bad.size == 0, f"{bad.size} birds differ, first {bad[:5]}"


def stacked_boundary_layout(O, n, side, d, seed):
    """Uniform birds where, in every group of four, the second and third sit
    exactly on the first (distance 0, the same cell) and the fourth sits on
    a cell corner or one ulp to either side of it."""
    x, y, hx, hy = O.init_uniform(n, side, side, seed, 1)
    d32, s32 = np.float32(d), np.float32(side)
    for i in range(n):
        r = i % 4
        if r in (1, 2):
            x[i], y[i] = x[i - r], y[i - r]
        elif r == 3:
            bx = np.float32(np.float32(int(x[i] / d32)) * d32)
            by = np.float32(np.float32(int(y[i] / d32)) * d32)
            m = (i // 4) % 3
            if m == 1:
                bx, by = np.nextafter(bx, np.float32(0)), np.nextafter(by, s32)
            elif m == 2:
                bx, by = np.nextafter(bx, s32), np.nextafter(by, np.float32(0))
            x[i] = bx if bx < s32 else np.float32(0)
            y[i] = by if by < s32 else np.float32(0)
    return x, y, hx, hy


@pytest.mark.parametrize("binned", [False, True])
def test_stacked_birds_and_cell_boundaries_match_oracle(oracle, binned):
    """Birds at identical positions (squared distance 0 in Bird::step's
    avoidance term, ties in the bag order broken by id) and on cell
    corners, where discretize and check_circle meet their boundaries: the
    neighbour lists and 15 steps equal the oracle's."""
    O = oracle
    n, side, steps = 3000, 200.0, 15
    x, y, hx, hy = stacked_boundary_layout(O, n, side, D, 17)
    assert np.unique(np.stack([x, y]), axis=1).shape[1] < n * 3 // 4 + 10
    q = Field2D(side, side, D, True, n)
    q.set_object_locations(np.arange(n, dtype=np.uint32), x, y, hx, hy)
    q.lazy_update()
    of = O.Field2D(side, side, D, True)
    for i in range(n):
        of.set_object_location(i, float(x[i]), float(y[i]))
    of.lazy_update()
    storage, row_ptr, ids = q.neighbor_sets(10.0)
    a = q.read_agents(0)
    for k in range(0, len(storage), 7):
        want = [b[0] for b in of.get_neighbors_within_distance(float(a["x"][k]), float(a["y"][k]),
                                                               10.0)]
        assert list(ids[row_ptr[k]:row_ptr[k + 1]]) == want, (k, storage[k])
    q.close()
    f = Field2D(side, side, D, True, n)
    f.set_object_locations(np.arange(n, dtype=np.uint32), x, y, hx, hy)
    f.schedule_run(FlockParams(seed=17), 0, steps, binned=binned)
    st, got_ids = f.state_by_id(0)
    f.close()
    fl = O.Flocker(side, side, D, True, x, y, hx, hy, O.params(seed=17))
    fl.step(steps)
    assert np.array_equal(got_ids, np.arange(n, dtype=np.uint32))
    ref = fl.state()
    bad = np.flatnonzero(np.any(st.view(np.uint32) != ref.view(np.uint32), axis=0))
    assert bad.size == 0, f"{bad.size} birds differ, first {bad[:5]}"
