"""Field2D on the GPU: the reference's own Field2D tests, ported.

tests/engine/field_2d.rs:30-247 (sequential Field2D) run against the device
Field2D through the C ABI.  The flockers fixture there is
Flocker::new((10, 10), n) with discretization 0.5, toroidal
(tests/model/flockers/state.rs:11-14).  Where the reference draws with
rand::rng(), these tests sweep the values instead (fly in 5..10) or use the
seeded Philox init.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = H = 10.0
D = 0.5


def make(capacity=16):
    from krabmaga_amd import Field2D
    return Field2D(W, H, D, True, capacity)


def test_field_2d_single_step():
    # tests/engine/field_2d.rs:30-50
    from krabmaga_amd import FlockParams
    f = make(10)
    f.init_uniform(10, seed=7, heading_mode=0)
    f.schedule_run(FlockParams(radius=10.0, seed=7), 0, 1)
    assert f.nagents == 10
    assert len(f.get_neighbors_within_distance(5.0, 5.0, 10.0)) == 10
    assert len(f.get_neighbors_within_relax_distance(5.0, 5.0, 10.0)) == 10


@pytest.mark.parametrize("fly", [5, 6, 7, 8, 9])
def test_field_2d_neighbors(fly):
    # tests/engine/field_2d.rs:57-117
    f = make()
    f.set_object_location(1, 0.0, 0.0)
    f.set_object_location(2, 0.0, 0.0)
    f.lazy_update()
    assert f.nagents == 2
    assert len(f.get_neighbors_within_distance(5.0, 5.0, 1.0)) == 0
    assert len(f.get_neighbors_within_relax_distance(5.0, 5.0, 1.0)) == 0
    b1 = (float(fly), float(fly))
    f.set_object_location(1, *b1)
    f.set_object_location(2, 0.0, 0.0)
    f.lazy_update()
    assert list(f.get_neighbors_within_distance(b1[0], b1[1], 1.0)) == [1]
    assert list(f.get_neighbors_within_distance(0.0, 0.0, 1.0)) == [2]
    assert sorted(f.get_neighbors_within_distance(5.0, 5.0, 10.0)) == [1, 2]
    assert sorted(f.get_neighbors_within_relax_distance(5.0, 5.0, 10.0)) == [1, 2]


def test_field_2d_gets():
    # tests/engine/field_2d.rs:124-178
    f = make()
    f.set_object_location(1, 0.0, 0.0)
    f.set_object_location(2, 5.0, 5.0)
    f.set_object_location(3, 5.0, 5.0)
    f.lazy_update()
    assert f.nagents == 3
    assert sorted(f.get_objects(5.0, 5.0)) == [2, 3]
    assert len(f.get_objects(10.0, 0.0)) == 0
    assert f.num_objects_at_location(5.0, 5.0) == 2
    assert f.num_objects_at_location(0.0, 0.0) == 1
    f.set_object_location(4, 0.0, 0.0)
    assert list(f.get_objects_unbuffered(0.0, 0.0)) == [4]
    assert list(f.get_objects(0.0, 0.0)) == [1]
    f.remove_object_location(4, 0.0, 0.0)
    assert len(f.get_objects_unbuffered(0.0, 0.0)) == 0


def test_field_2d_bags():
    # tests/engine/field_2d.rs:185-210 (dh = dw = ceil(10/0.5) + 1 = 21)
    f = make()
    assert f.num_empty_bags() == 21 * 21
    f.set_object_location(1, 0.0, 0.0)
    f.set_object_location(2, 0.0, 0.0)
    f.set_object_location(3, 4.0, 4.0)
    f.lazy_update()
    assert f.num_empty_bags() == 21 * 21 - 2


def test_field_2d_iter():
    # tests/engine/field_2d.rs:217-247
    f = make()
    locs = {1: (0.0, 0.0), 2: (0.01, 0.01), 3: (5.0, 5.0)}
    for i, (x, y) in locs.items():
        f.set_object_location(i, x, y)
    wb = f.read_agents(1)
    for i, x, y in zip(wb["id"], wb["x"], wb["y"]):
        assert i in f.get_objects_unbuffered(x, y)
    f.lazy_update()
    assert len(f.get_objects(0.0, 0.0)) == 2
    rb = f.read_agents(0)
    assert sorted(rb["id"]) == [1, 2, 3]
    for i, x, y in zip(rb["id"], rb["x"], rb["y"]):
        assert i in f.get_objects(x, y)


def test_write_buffer_dropped_by_lazy_update():
    # field_2d.rs:905-921: after a swap the new write buffer is empty
    f = make()
    f.set_object_location(1, 1.0, 1.0)
    f.lazy_update()
    assert f.count(1) == 0
    f.lazy_update()
    assert f.count(0) == 0


def test_out_of_grid_location_is_an_error():
    # the reference panics on an out-of-range bag index (field_2d.rs:840-842)
    from krabmaga_amd import KgError
    f = make()
    with pytest.raises(KgError):
        f.set_object_location(1, 20.0, 1.0)
    with pytest.raises(KgError):
        f.set_object_location(1, -1.0, 1.0)


@pytest.mark.parametrize("binned", [False, True])
def test_empty_flock_steps(binned):
    # Flocker::new(dim, 0): init places no bird, so every Schedule::step has
    # nothing to run and the field stays empty
    from krabmaga_amd import Field2D, FlockParams
    f = Field2D(200.0, 200.0, 10.0 / 1.5, True, 1)
    f.init_uniform(0, seed=3)
    f.schedule_run(FlockParams(seed=3), 0, 6, binned=binned)
    f.sync()
    assert f.count(0) == 0 and f.count(1) == 0
    assert f.get_neighbors_within_distance(100.0, 100.0, 10.0).size == 0
    assert f.get_neighbors_within_relax_distance(0.0, 0.0, 10.0).size == 0
    dw = int(np.ceil(np.float32(200.0) / np.float32(10.0 / 1.5))) + 1
    assert f.num_empty_bags() == dw * dw
    st, ids = f.state_by_id(0)
    assert st.shape == (4, 0) and ids.size == 0
    f.close()
