"""Field2D on the device, with the reference's method names.

Mirrors src/engine/fields/field_2d.rs (sequential Field2D, :295-921) over the
C ABI of libkrabgpu.so.  Agents are flockers birds: an id, a position and a
heading (Bird{id, pos, last_d}, tests/model/flockers/bird.rs:19-25).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import check, f32p, u32p


@dataclass
class FlockParams:
    """Weights of Bird::step (bird.rs:12-17), radius (bird.rs:41), Philox seed."""
    cohesion: float = 1.0
    avoidance: float = 1.0
    consistency: float = 1.0
    randomness: float = 1.0
    momentum: float = 1.0
    jump: float = 0.7
    radius: float = 10.0
    seed: int = 0

    def c(self):
        return _abi.FlockParams(self.cohesion, self.avoidance, self.consistency, self.randomness,
                                self.momentum, self.jump, self.radius, 0, self.seed)


class Field2D:
    """Field2D::new(w, h, d, toroidal) living in HBM (field_2d.rs:304-322)."""

    def __init__(self, width, height, discretization, toroidal, capacity, bin_rows=0,
                 tile_cols=0, tile_rows=0):
        self._L = _abi.lib()
        self.width = float(width)
        self.height = float(height)
        self.discretization = float(discretization)
        self.toroidal = bool(toroidal)
        self.capacity = int(capacity)
        desc = _abi.FieldDesc(width, height, discretization, int(bool(toroidal)), capacity, bin_rows,
                              tile_cols, tile_rows)
        h = ctypes.c_void_p()
        check(self._L.kg_field_create(ctypes.byref(desc), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.kg_field_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- field_2d.rs:838-846 -------------------------------------------------
    def set_object_location(self, bird_id, x, y, hx=0.0, hy=0.0):
        self.set_object_locations([bird_id], [x], [y], [hx], [hy])

    def set_object_locations(self, ids, x, y, hx=None, hy=None):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        hx = None if hx is None else np.ascontiguousarray(hx, dtype=np.float32)
        hy = None if hy is None else np.ascontiguousarray(hy, dtype=np.float32)
        check(self._L.kg_set_object_locations(self._h, u32p(ids), f32p(x), f32p(y), f32p(hx),
                                              f32p(hy), len(ids)))

    def init_uniform(self, n, seed, heading_mode=1):
        """Flocker::init on the device (state.rs:41-56, Philox draws)."""
        check(self._L.kg_init_uniform(self._h, n, seed, heading_mode))

    # -- field_2d.rs:885-898 -------------------------------------------------
    def remove_object_location(self, bird_id, x, y):
        check(self._L.kg_remove_object_location(self._h, bird_id, x, y))

    # -- field_2d.rs:905-921 -------------------------------------------------
    def lazy_update(self):
        check(self._L.kg_lazy_update(self._h))

    def _ids_call(self, fn, *args):
        cap = 256
        while True:
            out = np.zeros(cap, dtype=np.uint32)
            n = ctypes.c_size_t()
            rc = fn(self._h, *args, u32p(out), cap, ctypes.byref(n))
            if rc == _abi.KG_E_CAPACITY:
                cap = int(n.value)
                continue
            check(rc)
            return out[: n.value].copy()

    # -- field_2d.rs:386-516 -------------------------------------------------
    def get_neighbors_within_distance(self, x, y, dist):
        return self._ids_call(self._L.kg_get_neighbors_within_distance, x, y, dist)

    def get_neighbors_within_relax_distance(self, x, y, dist):
        return self._ids_call(self._L.kg_get_neighbors_within_relax_distance, x, y, dist)

    # -- field_2d.rs:544-582, 803-808, 727-739 -------------------------------
    def get_objects(self, x, y):
        return self._ids_call(self._L.kg_get_objects, x, y, 0)

    def get_objects_unbuffered(self, x, y):
        return self._ids_call(self._L.kg_get_objects, x, y, 1)

    def num_objects_at_location(self, x, y):
        n = ctypes.c_size_t()
        check(self._L.kg_num_objects_at_location(self._h, x, y, ctypes.byref(n)))
        return n.value

    def num_empty_bags(self):
        n = ctypes.c_size_t()
        check(self._L.kg_num_empty_bags(self._h, ctypes.byref(n)))
        return n.value

    # -- field_2d.rs:727-800 --------------------------------------------------
    def get_empty_bags(self):
        """not_discretize positions (x, y) of every empty bag of the read
        buffer, in bag index order (two float32 arrays)."""
        cap = 0
        while True:
            x = np.zeros(max(cap, 1), np.float32)
            y = np.zeros(max(cap, 1), np.float32)
            n = ctypes.c_size_t()
            rc = self._L.kg_get_empty_bags(self._h, f32p(x), f32p(y), cap, ctypes.byref(n))
            if rc == _abi.KG_E_CAPACITY:
                cap = int(n.value)
                continue
            check(rc)
            return x[: n.value].copy(), y[: n.value].copy()

    def get_random_empty_bag(self, seed):
        """A seeded draw among the empty bags: (x, y), or None if there is
        none (field_2d.rs:764-800)."""
        x, y, found = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
        check(self._L.kg_get_random_empty_bag(self._h, seed, ctypes.byref(x), ctypes.byref(y),
                                              ctypes.byref(found)))
        return (x.value, y.value) if found.value else None

    # -- field_2d.rs:619-726 ---------------------------------------------------
    def iter_objects(self, closure, unbuffered=False):
        """closure((bag_x, bag_y), bird) for every agent, bag by bag; bird is
        (id, x, y, hx, hy)."""
        def fn(_user, bx, by, i, x, y, hx, hy):
            closure((bx, by), (i, x, y, hx, hy))
        cb = _abi.OBJECT_FN(fn)
        check(self._L.kg_iter_objects(self._h, 1 if unbuffered else 0, cb, None))

    def iter_objects_unbuffered(self, closure):
        self.iter_objects(closure, unbuffered=True)

    @property
    def nagents(self):
        n = ctypes.c_size_t()
        check(self._L.kg_nagents(self._h, ctypes.byref(n)))
        return n.value

    def count(self, buffer=0):
        n = ctypes.c_size_t()
        check(self._L.kg_count(self._h, buffer, ctypes.byref(n)))
        return n.value

    def neighbor_sets(self, dist):
        """All-agent neighbour CSR over the read buffer's storage order.

        Returns (storage_ids, row_ptr, ids)."""
        n = self.count(0)
        row_ptr = np.zeros(n + 1, dtype=np.uint32)
        total = ctypes.c_size_t()
        cap = max(16, n * 4)
        while True:
            ids = np.zeros(cap, dtype=np.uint32)
            rc = self._L.kg_neighbor_sets(self._h, dist, u32p(row_ptr), u32p(ids), cap,
                                          ctypes.byref(total))
            if rc == _abi.KG_E_CAPACITY:
                cap = int(total.value)
                continue
            check(rc)
            break
        storage = self.read_agents(0)["id"]
        return storage, row_ptr, ids[: total.value].copy()

    # -- Bird::step for all agents / the simulate! loop ----------------------
    def flock_step(self, params, step):
        p = params.c()
        check(self._L.kg_flock_step(self._h, ctypes.byref(p), step))

    def schedule_run(self, params, first_step, nsteps, use_graph=True, binned=False,
                     tile_resident=False):
        """`nsteps` Schedule::step calls from `first_step` (simulate!'s loop).

        By default fields of fewer than 32,768 agents take the one-launch
        tile-resident step and larger ones the three-kernel path; binned=True
        forces the three-kernel path, tile_resident=True the tile-resident
        step where the field suits it (same results either way)."""
        p = params.c()
        flags = (1 if use_graph else 0) | (2 if binned else 0) | (4 if tile_resident else 0)
        check(self._L.kg_schedule_run(self._h, ctypes.byref(p), first_step, nsteps, flags))

    def read_agents(self, buffer=0):
        n = ctypes.c_size_t()
        rc = self._L.kg_read_agents(self._h, buffer, None, None, None, None, None, 0,
                                    ctypes.byref(n))
        if rc not in (_abi.KG_OK, _abi.KG_E_CAPACITY):
            check(rc)
        m = int(n.value)
        out = {"id": np.zeros(m, np.uint32), "x": np.zeros(m, np.float32),
               "y": np.zeros(m, np.float32), "hx": np.zeros(m, np.float32),
               "hy": np.zeros(m, np.float32)}
        check(self._L.kg_read_agents(self._h, buffer, u32p(out["id"]), f32p(out["x"]),
                                     f32p(out["y"]), f32p(out["hx"]), f32p(out["hy"]), m,
                                     ctypes.byref(n)))
        return out

    def state_by_id(self, buffer=0):
        """Read buffer as a (4, n) array [x, y, hx, hy] indexed by id 0..n-1."""
        a = self.read_agents(buffer)
        order = np.argsort(a["id"], kind="stable")
        return np.stack([a["x"][order], a["y"][order], a["hx"][order], a["hy"][order]]), a["id"][order]

    def profile(self, enable=True):
        """Time every kernel launch with HIP events (disables graph replay)."""
        check(self._L.kg_field_profile(self._h, int(bool(enable))))

    def profile_read(self, kernel):
        """(total_ms, launches) of one kernel since profile() was called."""
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        check(self._L.kg_field_profile_read(self._h, _abi.KERNELS[kernel], ctypes.byref(ms),
                                            ctypes.byref(n)))
        return ms.value, n.value

    STAT_NAMES = ("rt_active", "rt_unsuitable", "rt_plans", "rt_flags", "tiles", "tile_cols",
                  "tile_rows", "bin_rows", "lds_runs", "lds_own", "lds_extras", "lds_bytes",
                  "lds_limit", "max_tile_cap", "dense_tiles", "rt_steps")

    def stats(self):
        """Layout statistics (see kg_field_stats)."""
        out = (ctypes.c_uint64 * 16)()
        check(self._L.kg_field_stats(self._h, out))
        return dict(zip(self.STAT_NAMES, (int(v) for v in out)))

    def sync(self):
        check(self._L.kg_field_sync(self._h))

    @property
    def stream(self):
        return self._L.kg_field_stream(self._h)
