"""Model exploration of flockers over the device ensemble.

The reference sweeps a model's parameters with `gen_param!` (src/lib.rs:
1781-1816), combines them per `ExploreMode` (src/lib.rs:481-486;
`build_configurations!`, src/lib.rs:1724-1748), runs `n_conf * rep_conf`
simulations with `explore_sequential!` (src/explore/model_exploration.rs:
231-313) or `explore_parallel!` (:354-432), collects one `FrameRow` per run
(`build_dataframe!`, :451-540) and writes them with `write_csv`
(src/lib.rs:1762-1773); `load_csv!` (src/lib.rs:1878-1900) reads columns
back.  Here the parallel sweep admits every run into one Ensemble, so a
single launch per schedule step (per tile class) advances all of them; the
sequential sweep runs each on its own Field2D, one after another.  A run's
birds are the same bits either way.

Differences from the reference, each forced by the device engine:
- the reference's runs draw from an unseeded thread RNG; here run r of a
  sweep uses the Philox keys (seed << 32) | r, so a sweep is reproducible;
- in the parallel sweep, the runs of one batch share one timed region, so
  their "Run Duration" and "Step per sec" are the batch's (with an
  end_condition, a run's duration ends at the step that stopped it);
- `end_condition` (flockers' is the default `false`) is a host callable:
  given one, runs advance one schedule step at a time and it is consulted
  after each step, as simulate_explore! does (model_exploration.rs:
  173-177); without one, steps run in batches of device graphs.
"""
import csv
import enum
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import stats
from .ensemble import Ensemble, RunSpec, footprint
from .field import Field2D, FlockParams


class ExploreMode(enum.Enum):
    """src/lib.rs:481-486 (the reference's spelling `Exaustive` is kept)."""
    Exaustive = 0
    Matched = 1
    Exhaustive = 0  # an alias of Exaustive


def gen_param(kind, lo, hi, n=1, rng=None):
    """`gen_param!(type, min, max, n)` (src/lib.rs:1794-1816): n values drawn
    uniformly from [min, max).  Bounds are swapped when min > max and max is
    raised by one when they are equal.  The reference sets n = 1 when n is 0
    but then takes the original n values, so n = 0 yields an empty list."""
    if kind not in (int, float):
        raise TypeError("kind must be int or float")
    lo, hi = kind(lo), kind(hi)
    if lo > hi:
        lo, hi = hi, lo
    elif lo == hi:
        hi = hi + kind(1)
    n = int(n)
    rng = rng if rng is not None else np.random.default_rng()
    if kind is int:
        return [int(v) for v in rng.integers(lo, hi, size=n)]
    return [float(v) for v in rng.uniform(lo, hi, size=n)]


def build_configurations(sizes):
    """Index table of the exhaustive sweep (`build_configurations!`,
    src/lib.rs:1724-1748): row k gives, for every configuration, the index
    into input k.  Input k's index changes every n_conf / (|in_0|...|in_k|)
    configurations, so the first input varies slowest."""
    sizes = [int(s) for s in sizes]
    if any(s <= 0 for s in sizes):
        # the reference divides by the input's length here
        raise ValueError("every input of an exhaustive sweep needs at least one value")
    n_conf = math.prod(sizes)
    rows = []
    rep = n_conf
    for s in sizes:
        rep //= s
        rows.append([(i // rep) % s for i in range(n_conf)])
    return rows


def configurations(inputs, mode):
    """The parameter dicts of a sweep, in configuration order
    (model_exploration.rs:369-381 and :388-412)."""
    names = list(inputs)
    values = [list(inputs[k]) for k in names]
    if mode is ExploreMode.Exaustive:
        rows = build_configurations([len(v) for v in values])
        n_conf = len(rows[0]) if rows else 1
        return [{k: values[j][rows[j][i]] for j, k in enumerate(names)} for i in range(n_conf)]
    if mode is ExploreMode.Matched:
        # n_conf is the last input's length; every input is indexed with the
        # configuration number, and a shorter one is out of range
        n_conf = len(values[-1]) if values else 1
        for k, v in zip(names, values):
            if len(v) < n_conf:
                raise IndexError(f"matched sweep: input {k!r} has {len(v)} values, "
                                 f"{n_conf} configurations need one each")
        return [{k: v[i] for k, v in zip(names, values)} for i in range(n_conf)]
    raise ValueError(f"unknown explore mode {mode!r}")


# ---- the flockers State: inputs it takes, outputs it reports ----

D = 10.0 / 1.5

FLOCKERS_INPUTS = {
    # Flocker::new(dim, initial_flockers) (tests/model/flockers/state.rs:24-32)
    # at 1,000 birds on 200 x 200 with cells of 10 / 1.5, and the Bird::step
    # constants (tests/model/flockers/bird.rs:12-17, 41)
    "agents": 1000, "width": 200.0, "height": 200.0, "discretization": D,
    "cohesion": 1.0, "avoidance": 1.0, "consistency": 1.0, "randomness": 1.0,
    "momentum": 1.0, "jump": 0.7, "radius": 10.0, "heading_mode": 1,
}


def flockers_spec(conf, seed):
    """RunSpec of one configuration; unknown input names are an error."""
    bad = set(conf) - set(FLOCKERS_INPUTS)
    if bad:
        raise KeyError(f"flockers takes no input named {sorted(bad)}")
    c = dict(FLOCKERS_INPUTS, **conf)
    p = FlockParams(cohesion=float(c["cohesion"]), avoidance=float(c["avoidance"]),
                    consistency=float(c["consistency"]), randomness=float(c["randomness"]),
                    momentum=float(c["momentum"]), jump=float(c["jump"]),
                    radius=float(c["radius"]), seed=int(seed))
    return RunSpec(width=float(c["width"]), height=float(c["height"]), agents=int(c["agents"]),
                   params=p, discretization=float(c["discretization"]),
                   heading_mode=int(c["heading_mode"]), init_seed=int(seed))


def _polarisation(st, spec):
    return stats.polarisation(st[2], st[3])


def _mean_nn(st, spec):
    return stats.mean_nn_distance(st[0], st[1], spec.width, spec.height)


def _mean_x(st, spec):
    return float(np.mean(st[0], dtype=np.float64))


def _mean_y(st, spec):
    return float(np.mean(st[1], dtype=np.float64))


FLOCKERS_OUTPUTS = {"polarisation": _polarisation, "mean_nn": _mean_nn, "mean_x": _mean_x,
                    "mean_y": _mean_y}


# ---- FrameRow, write_csv, load_csv ----

def _rust_float(v, f32, debug):
    """Rust's `{:?}` (debug=True) or `to_string()` of an f32 / f64."""
    v = np.float32(v) if f32 else np.float64(v)
    if np.isnan(v):
        return "NaN"
    if np.isinf(v):
        return "inf" if v > 0 else "-inf"
    a = abs(float(v))
    if debug and a != 0.0 and (a < 1e-4 or a >= 1e16):
        m, e = np.format_float_scientific(v, unique=True, trim="-").split("e")
        return f"{m}e{int(e)}"
    s = np.format_float_positional(v, unique=True, trim="-")
    if debug and "." not in s:
        s += ".0"
    return s


def rust_debug(v):
    """`format!("{:?}", v)` for the value types a FrameRow column holds."""
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, np.float32):
        return _rust_float(v, True, True)
    if isinstance(v, (float, np.floating)):
        return _rust_float(v, False, True)
    if isinstance(v, (list, tuple, np.ndarray)):
        return "[" + ", ".join(rust_debug(x) for x in v) + "]"
    return str(v)


@dataclass
class FrameRow:
    """One run of a sweep (`build_dataframe!`, model_exploration.rs:451-540)."""
    conf_num: int
    conf_rep: int
    inputs: dict
    outputs: dict
    run_duration: float
    step_per_sec: float
    state: object = field(default=None, repr=False)

    @staticmethod
    def field_names(input_names, output_names):
        return ["Simulation", "Run", *input_names, *output_names, "Run Duration", "Step per sec"]

    def to_strings(self):
        return ([str(self.conf_num), str(self.conf_rep)]
                + [rust_debug(v) for v in self.inputs.values()]
                + [rust_debug(v) for v in self.outputs.values()]
                + [_rust_float(self.run_duration, True, False),
                   _rust_float(self.step_per_sec, True, False)])


def write_csv(name, rows):
    """`write_csv(name, dataframe)` (src/lib.rs:1762-1773): `<name>.csv`, a
    header row, then one record per run.  An empty dataframe writes nothing,
    as the reference's callers skip it (`if !result.is_empty()`)."""
    rows = list(rows)
    if not rows:
        raise ValueError("empty dataframe")
    with open(f"{name}.csv", "w", newline="") as f:
        w = csv.writer(f, lineterminator="\n")
        w.writerow(FrameRow.field_names(list(rows[0].inputs), list(rows[0].outputs)))
        for r in rows:
            w.writerow(r.to_strings())


def load_csv(path, *columns):
    """`load_csv!(file, x: T, ...)` (src/lib.rs:1878-1900): the header row is
    skipped and column i of every record is parsed as the i-th type.
    `columns` are (name, type) pairs; returns one list per column."""
    out = [[] for _ in columns]
    with open(path, newline="") as f:
        r = csv.reader(f)
        next(r, None)
        for rec in r:
            for i, (_, kind) in enumerate(columns):
                out[i].append(_parse(rec[i], kind))
    return tuple(out)


def _parse(s, kind):
    if kind is bool:
        if s not in ("true", "false"):
            raise ValueError(f"not a bool: {s!r}")
        return s == "true"
    if kind is int:
        return int(s)
    if kind is float:
        if s in ("NaN", "inf", "-inf"):
            return float(s.replace("NaN", "nan"))
        return float(s)
    return kind(s)


# ---- the sweep ----

def explore(nstep, rep_conf, inputs, outputs=("polarisation",), mode=ExploreMode.Matched,
            parallel=True, seed=0, hbm_budget=200 << 30, tuning=None, keep_state=False,
            end_condition=None):
    """Runs every configuration `rep_conf` times for `nstep` schedule steps
    and returns one FrameRow per run, configurations in order, repetitions
    in order within each.

    inputs: {name: list of values} for the flockers State (FLOCKERS_INPUTS
    names the accepted ones and their defaults).  outputs: names from
    FLOCKERS_OUTPUTS, or {name: fn(state, spec)} where state is the (4, n)
    [x, y, hx, hy] array by bird id.  parallel=True is explore_parallel!
    (one Ensemble; repetitions numbered from 0), False explore_sequential!
    (one Field2D per run, one after another; repetitions numbered from 1).
    Parallel batches are cut so that their device footprint stays within
    hbm_budget bytes.

    end_condition(step, state, spec) -> bool, if given, is called after
    every schedule step of every run (State::end_condition after
    schedule.step, model_exploration.rs:173-177): step is the number of
    steps the run has taken, state() reads its (4, n) array by bird id.  A
    run that returns True stops there; its "Step per sec" counts the steps
    it took, as the reference divides schedule.step by the duration."""
    if nstep < 0 or rep_conf < 0:
        raise ValueError("nstep and rep_conf must be non-negative")
    if isinstance(outputs, str):
        outputs = (outputs,)
    outs = ({k: FLOCKERS_OUTPUTS[k] for k in outputs} if not isinstance(outputs, dict)
            else dict(outputs))
    confs = configurations(inputs, mode)
    runs = [(i, j) for i in range(len(confs)) for j in range(rep_conf)]
    specs = [flockers_spec(confs[i], (int(seed) << 32) | (i * rep_conf + j)) for i, j in runs]
    rows = [None] * len(runs)

    def finish(k, st, dt, steps=nstep):
        i, j = runs[k]
        vals = {name: fn(st, specs[k]) for name, fn in outs.items()}
        dt32 = np.float32(dt)
        with np.errstate(divide="ignore", invalid="ignore"):
            sps = np.float32(steps) / dt32  # f32 division as in Rust: x / 0 = inf, 0 / 0 = NaN
        rows[k] = FrameRow(i, j if parallel else j + 1, dict(confs[i]), vals, float(dt32),
                           float(sps), st if keep_state else None)

    if not parallel:
        for k, s in enumerate(specs):
            f = Field2D(s.width, s.height, s.discretization, True, s.agents)
            try:
                f.init_uniform(s.agents, s.init_seed, s.heading_mode)
                f.sync()
                t0 = time.perf_counter()
                done = nstep
                if end_condition is None:
                    f.schedule_run(s.params, 0, nstep)
                else:
                    for step in range(nstep):
                        f.schedule_run(s.params, step, 1)
                        if end_condition(step + 1, lambda: f.state_by_id(0)[0], s):
                            done = step + 1
                            break
                f.sync()
                dt = time.perf_counter() - t0
                # before its first lazy_update the birds are in the write buffer
                st, _ = f.state_by_id(0 if done > 0 else 1)
            finally:
                f.close()
            finish(k, st, dt, done)
        return rows

    for batch in _batches([footprint(s, tuning) for s in specs], hbm_budget):
        with Ensemble(tuning) as ens:
            ids = ens.admit([specs[k] for k in batch])
            ens.sync()
            t0 = time.perf_counter()
            if end_condition is None:
                ens.run(nstep)
                ens.sync()
                dt = time.perf_counter() - t0
                for k, rid in zip(batch, ids):
                    st, _ = ens.state_by_id(rid)
                    finish(k, st, dt)
                continue
            live = dict(zip(batch, ids))
            for step in range(nstep):
                ens.run(1)
                ens.sync()
                for k, rid in list(live.items()):
                    if end_condition(step + 1, lambda rid=rid: ens.state_by_id(rid)[0], specs[k]):
                        dt = time.perf_counter() - t0
                        st, _ = ens.state_by_id(rid)
                        finish(k, st, dt, step + 1)
                        ens.retire(rid)
                        del live[k]
                if not live:
                    break
            ens.sync()
            dt = time.perf_counter() - t0
            for k, rid in live.items():
                st, _ = ens.state_by_id(rid)
                finish(k, st, dt)
    return rows


def _batches(sizes, budget):
    """Consecutive groups of run indices whose byte sizes sum to at most
    budget (a run larger than the budget is an error)."""
    out, cur, used = [], [], 0
    for k, b in enumerate(sizes):
        if b > budget:
            raise ValueError(f"run {k} needs {b} bytes, more than the budget of {budget}")
        if cur and used + b > budget:
            out.append(cur)
            cur, used = [], 0
        cur.append(k)
        used += b
    if cur:
        out.append(cur)
    return out


__all__ = ["ExploreMode", "gen_param", "build_configurations", "configurations", "explore",
           "FrameRow", "write_csv", "load_csv", "rust_debug", "flockers_spec", "FLOCKERS_INPUTS",
           "FLOCKERS_OUTPUTS"]

