"""Ensembles of flockers runs on the device (explore_parallel!'s workload).

src/explore/model_exploration.rs:354-432 runs `n_conf * rep_conf`
independent simulations, one per CPU core at a time.  An Ensemble advances
all of them together on the GPU: every run keeps its own Field2D geometry,
flock parameters, Philox key and schedule step, and one kernel launch per
schedule step (per tile class) moves every run.  A run's birds are the bits
it would have alone (Field2D.schedule_run on its own field).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import check, f32p, u32p
from .field import FlockParams


@dataclass
class RunSpec:
    """One run: Field2D::new(width, height, discretization, true) with
    `agents` birds from Flocker::init (Philox key `init_seed`)."""
    width: float
    height: float
    agents: int
    params: FlockParams = field(default_factory=FlockParams)
    discretization: float = 10.0 / 1.5
    heading_mode: int = 1
    init_seed: int = 0

    def c(self):
        return _abi.RunDesc(self.width, self.height, self.discretization, 1, self.agents,
                            self.heading_mode, self.init_seed, self.params.c())


@dataclass
class Tuning:
    """Layout choices (0 = automatic): tile size in cells, bin height, LDS
    bytes per normal-class work-group."""
    tile_cols: int = 0
    tile_rows: int = 0
    bin_rows: int = 0
    lds_normal: int = 0

    def c(self):
        return _abi.Tuning(self.tile_cols, self.tile_rows, self.bin_rows, self.lds_normal)


def footprint(spec, tuning=None):
    """Device bytes `spec` occupies in an ensemble."""
    b = ctypes.c_uint64()
    t = (tuning or Tuning()).c()
    check(_abi.lib().kg_run_footprint(ctypes.byref(spec.c()), ctypes.byref(t), ctypes.byref(b)))
    return b.value


class Ensemble:
    STAT_NAMES = ("runs", "agents", "tiles", "dense_tiles", "builds", "flags", "steps",
                  "bytes")

    def __init__(self, tuning=None):
        self._L = _abi.lib()
        h = ctypes.c_void_p()
        t = (tuning or Tuning()).c()
        check(self._L.kg_ensemble_create(ctypes.byref(t), ctypes.byref(h)))
        self._h = h
        self._agents = {}

    def close(self):
        if getattr(self, "_h", None):
            self._L.kg_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def admit(self, specs):
        """Adds runs; returns their ids (step 0 comes with the next run())."""
        specs = list(specs)
        arr = (_abi.RunDesc * len(specs))(*[s.c() for s in specs])
        ids = np.zeros(len(specs), np.uint32)
        check(self._L.kg_ensemble_admit(self._h, arr, len(specs), u32p(ids)))
        for i, s in zip(ids, specs):
            self._agents[int(i)] = s.agents
        return [int(i) for i in ids]

    def retire(self, run_id):
        check(self._L.kg_ensemble_retire(self._h, run_id))
        self._agents.pop(run_id, None)

    def run(self, nsteps, use_graph=True, binned=False, tile_resident=False):
        """Every run advances nsteps schedule steps (asynchronous).

        By default ensembles whose runs average 32,768 agents or more step
        the bin-grouped layout with the three-kernel path, smaller ones the
        tile-resident layout; binned / tile_resident choose either (same
        results)."""
        flags = (1 if use_graph else 0) | (2 if binned else 0) | (4 if tile_resident else 0)
        check(self._L.kg_ensemble_run(self._h, nsteps, flags))

    def steps(self, run_id):
        s = ctypes.c_uint64()
        check(self._L.kg_ensemble_steps(self._h, run_id, ctypes.byref(s)))
        return s.value

    def read(self, run_id):
        m = self._agents[run_id]
        out = {"id": np.zeros(m, np.uint32), "x": np.zeros(m, np.float32),
               "y": np.zeros(m, np.float32), "hx": np.zeros(m, np.float32),
               "hy": np.zeros(m, np.float32)}
        n = ctypes.c_size_t()
        check(self._L.kg_ensemble_read(self._h, run_id, u32p(out["id"]), f32p(out["x"]),
                                       f32p(out["y"]), f32p(out["hx"]), f32p(out["hy"]), m,
                                       ctypes.byref(n)))
        return out

    def state_by_id(self, run_id):
        """(4, n) array [x, y, hx, hy] indexed by bird id, and the ids."""
        a = self.read(run_id)
        order = np.argsort(a["id"], kind="stable")
        return np.stack([a["x"][order], a["y"][order], a["hx"][order], a["hy"][order]]), a["id"][order]

    def sync(self):
        check(self._L.kg_ensemble_sync(self._h))

    @property
    def stream(self):
        return self._L.kg_ensemble_stream(self._h)

    def stats(self):
        out = (ctypes.c_uint64 * 8)()
        check(self._L.kg_ensemble_stats(self._h, out))
        return dict(zip(self.STAT_NAMES, (int(v) for v in out)))
