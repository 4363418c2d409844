"""Model-exploration front end (krabmaga_amd.explore) on the CPU: sweep
combinations, gen_param!'s range rules, the FrameRow text of each value
type, and write_csv / load_csv round trips.

The exhaustive sweep is pinned by itertools.product (the reference's index
table, src/lib.rs:1724-1748, lists the Cartesian product with the first
input slowest); the value formatting by the documented output of Rust's
`{:?}` and `to_string()` for integers, f32, f64 and bool."""
import itertools

import numpy as np
import pytest

from krabmaga_amd.explore import (ExploreMode, FrameRow, _batches, _rust_float,
                                  build_configurations, configurations, flockers_spec,
                                  gen_param, load_csv, rust_debug, write_csv)


@pytest.mark.parametrize("sizes", [[1], [3], [2, 3], [3, 2], [2, 1, 4], [1, 1, 1], [4, 3, 2, 2]])
def test_exhaustive_is_the_cartesian_product(sizes):
    inputs = {f"p{k}": [f"v{k}_{j}" for j in range(s)] for k, s in enumerate(sizes)}
    got = configurations(inputs, ExploreMode.Exaustive)
    want = [dict(zip(inputs, c)) for c in itertools.product(*inputs.values())]
    assert got == want
    rows = build_configurations(sizes)
    assert len(rows) == len(sizes) and all(len(r) == int(np.prod(sizes)) for r in rows)


def test_exhaustive_rejects_an_empty_input():
    with pytest.raises(ValueError):
        configurations({"a": [1, 2], "b": []}, ExploreMode.Exaustive)


def test_exhaustive_alias():
    assert ExploreMode.Exhaustive is ExploreMode.Exaustive


def test_matched_takes_the_last_inputs_length():
    # n_conf is the last input's length (model_exploration.rs:379-381); a
    # longer input's extra values are not used
    got = configurations({"a": [1, 2, 3, 4], "b": [10, 20]}, ExploreMode.Matched)
    assert got == [{"a": 1, "b": 10}, {"a": 2, "b": 20}]
    # a shorter input would be indexed out of range
    with pytest.raises(IndexError):
        configurations({"a": [1], "b": [10, 20]}, ExploreMode.Matched)


def test_gen_param_ranges():
    rng = np.random.default_rng(5)
    v = gen_param(int, 0, 10, 1000, rng)
    assert len(v) == 1000 and min(v) >= 0 and max(v) <= 9 and len(set(v)) == 10
    assert all(type(x) is int for x in v)
    f = gen_param(float, 10.0, 0.0, 1000, rng)   # bounds swapped
    assert min(f) >= 0.0 and max(f) < 10.0
    assert gen_param(int, 7, 7, 50, rng) == [7] * 50   # [7, 8)
    e = gen_param(float, 2.0, 2.0, 100, rng)
    assert all(2.0 <= x < 3.0 for x in e)
    assert gen_param(int, 0, 10, 0, rng) == []   # take(0) of the stream
    assert len(gen_param(float, 0, 1)) == 1
    assert gen_param(int, 0, 100, 20, np.random.default_rng(1)) == \
        gen_param(int, 0, 100, 20, np.random.default_rng(1))
    with pytest.raises(TypeError):
        gen_param(str, 0, 1)


@pytest.mark.parametrize("v,text", [
    (7, "7"), (np.uint32(4000000000), "4000000000"), (-3, "-3"), (True, "true"),
    (False, "false"), (1.0, "1.0"), (0.1, "0.1"), (0.0, "0.0"), (-2.5, "-2.5"),
    (1e16, "1e16"), (1.5e-7, "1.5e-7"), (1e-4, "0.0001"), (123456.789, "123456.789"),
    (9.99e15, "9990000000000000.0"), (-2.5e20, "-2.5e20"), (float("nan"), "NaN"),
    (float("inf"), "inf"), (np.float32(0.1), "0.1"), (np.float32(1.0), "1.0"),
    (np.float32(16777217.0), "16777216.0"), ([1, 2], "[1, 2]"), ([0.5, 2.0], "[0.5, 2.0]"),
])
def test_debug_text(v, text):
    assert rust_debug(v) == text


@pytest.mark.parametrize("v,text", [(1.0, "1"), (0.1, "0.1"), (3.25, "3.25"),
                                    (1e20, "100000000000000000000"), (1234.5678, "1234.5677"),
                                    (0.0, "0")])
def test_f32_display_text(v, text):
    assert _rust_float(v, True, False) == text


def test_csv_round_trip(tmp_path):
    rows = [FrameRow(i, j, {"agents": 100 * (i + 1), "jump": 0.25 * (i + 1)},
                     {"polarisation": 0.5 + 0.125 * j}, 0.5 + i, 1000.0 * (j + 1))
            for i in range(3) for j in range(2)]
    name = str(tmp_path / "explore_result")
    write_csv(name, rows)
    lines = open(name + ".csv").read().splitlines()
    assert lines[0] == "Simulation,Run,agents,jump,polarisation,Run Duration,Step per sec"
    assert lines[1] == "0,0,100,0.25,0.5,0.5,1000"
    assert len(lines) == 7
    conf, rep, agents, jump, pol, dur, sps = load_csv(
        name + ".csv", ("conf", int), ("rep", int), ("agents", int), ("jump", float),
        ("polarisation", float), ("dur", float), ("sps", float))
    assert conf == [r.conf_num for r in rows] and rep == [r.conf_rep for r in rows]
    assert agents == [r.inputs["agents"] for r in rows]
    assert jump == [r.inputs["jump"] for r in rows]
    assert pol == [r.outputs["polarisation"] for r in rows]
    assert dur == [r.run_duration for r in rows] and sps == [r.step_per_sec for r in rows]
    with pytest.raises(ValueError):
        write_csv(name, [])


def test_csv_quotes_fields_that_need_it(tmp_path):
    rows = [FrameRow(0, 0, {"w": [1.0, 2.0]}, {}, 1.0, 2.0)]
    name = str(tmp_path / "q")
    write_csv(name, rows)
    assert open(name + ".csv").read().splitlines()[1] == '0,0,"[1.0, 2.0]",1,2'


def test_batches_respect_the_budget():
    assert _batches([3, 3, 3, 5, 1], 6) == [[0, 1], [2], [3, 4]]
    assert _batches([], 6) == []
    with pytest.raises(ValueError):
        _batches([7], 6)


def test_flockers_spec_defaults_and_names():
    s = flockers_spec({"agents": 500, "jump": 0.4}, 9)
    assert (s.agents, s.width, s.height, s.params.jump, s.params.seed, s.init_seed) == \
        (500, 200.0, 200.0, 0.4, 9, 9)
    assert abs(s.discretization - 10.0 / 1.5) < 1e-12
    with pytest.raises(KeyError):
        flockers_spec({"agent": 5}, 0)


def test_sweep_without_inputs_is_one_default_configuration():
    # n_conf starts at 1 and no input changes it (model_exploration.rs:369-381)
    assert configurations({}, ExploreMode.Exaustive) == [{}]
    assert configurations({}, ExploreMode.Matched) == [{}]
