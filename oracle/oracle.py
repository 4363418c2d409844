"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle (oracle.c, oracle_par.c) restates krABMaga's sequential Field2D
(src/engine/fields/field_2d.rs:267-1014), the in-tree flockers model
(tests/model/flockers/bird.rs:38-154, state.rs:41-60) and Schedule::step
(src/engine/schedule.rs:347-413).  Only the test suite, the smoke check and
the CPU-baseline leg of bench.py may import this module; the product library
never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class Bird(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("px", ctypes.c_float), ("py", ctypes.c_float),
                ("hx", ctypes.c_float), ("hy", ctypes.c_float)]


class FlockParams(ctypes.Structure):
    _fields_ = [("cohesion", ctypes.c_float), ("avoidance", ctypes.c_float),
                ("consistency", ctypes.c_float), ("randomness", ctypes.c_float),
                ("momentum", ctypes.c_float), ("jump", ctypes.c_float),
                ("radius", ctypes.c_float), ("seed", ctypes.c_uint64)]


class RunSpec(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("side", ctypes.c_float),
                ("discretization", ctypes.c_float), ("steps", ctypes.c_uint32),
                ("p", FlockParams), ("heading_mode", ctypes.c_int)]


def build():
    """Compile liboracle.so with its Makefile (gcc, a second or two)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    L.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    L.or_u01.argtypes = [ctypes.c_uint32]
    L.or_u01.restype = ctypes.c_float
    L.or_step_uniforms.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, _f32p, _f32p]
    L.or_t_transform.argtypes = [ctypes.c_int32, ctypes.c_int32]
    L.or_t_transform.restype = ctypes.c_int32
    L.or_toroidal_transform.argtypes = [ctypes.c_float, ctypes.c_float]
    L.or_toroidal_transform.restype = ctypes.c_float
    L.or_toroidal_distance.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L.or_toroidal_distance.restype = ctypes.c_float
    L.or_distance.argtypes = [ctypes.c_float] * 6 + [ctypes.c_int]
    L.or_distance.restype = ctypes.c_float
    L.or_check_circle.argtypes = [ctypes.c_int32, ctypes.c_int32] + [ctypes.c_float] * 6 + [ctypes.c_int]
    L.or_check_circle.restype = ctypes.c_int
    L.or_f32_to_i32.argtypes = [ctypes.c_float]
    L.or_f32_to_i32.restype = ctypes.c_int32
    L.or_field_new.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    L.or_field_new.restype = P
    L.or_field_free.argtypes = [P]
    L.or_set_object_location.argtypes = [P, Bird, ctypes.c_float, ctypes.c_float]
    L.or_set_object_location.restype = ctypes.c_int
    L.or_remove_object_location.argtypes = [P, ctypes.c_uint32, ctypes.c_float, ctypes.c_float]
    L.or_lazy_update.argtypes = [P]
    for name in ("or_get_neighbors_within_distance", "or_get_neighbors_within_relax_distance"):
        fn = getattr(L, name)
        fn.argtypes = [P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.POINTER(Bird), ctypes.c_size_t]
        fn.restype = ctypes.c_size_t
    for name in ("or_get_objects", "or_get_objects_unbuffered"):
        fn = getattr(L, name)
        fn.argtypes = [P, ctypes.c_float, ctypes.c_float, ctypes.POINTER(Bird), ctypes.c_size_t]
        fn.restype = ctypes.c_size_t
    L.or_set_object_locations.argtypes = [P, _u32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_size_t]
    L.or_set_object_locations.restype = ctypes.c_int64
    L.or_neighbor_sets.argtypes = [P, _f32p, _f32p, ctypes.c_size_t, ctypes.c_float,
                                   ctypes.POINTER(ctypes.c_uint64), _u32p, ctypes.c_size_t]
    L.or_neighbor_sets.restype = ctypes.c_size_t
    L.or_num_objects_at_location.argtypes = [P, ctypes.c_float, ctypes.c_float]
    L.or_num_objects_at_location.restype = ctypes.c_size_t
    L.or_get_empty_bags.argtypes = [P, _f32p, _f32p, ctypes.c_size_t]
    L.or_get_empty_bags.restype = ctypes.c_size_t
    L.or_num_empty_bags.argtypes = [P]
    L.or_num_empty_bags.restype = ctypes.c_size_t
    L.or_nagents.argtypes = [P]
    L.or_nagents.restype = ctypes.c_size_t
    L.or_flocker_new.argtypes = [ctypes.c_uint32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_int, _f32p, _f32p, _f32p, _f32p, ctypes.POINTER(FlockParams)]
    L.or_flocker_new.restype = P
    L.or_flocker_free.argtypes = [P]
    L.or_flocker_set_order.argtypes = [P, _u32p]
    L.or_flocker_step.argtypes = [P]
    L.or_bird_step.argtypes = [P, ctypes.c_uint32]
    L.or_flocker_get.argtypes = [P, _f32p, _f32p, _f32p, _f32p]
    L.or_init_uniform.argtypes = [ctypes.c_uint32, ctypes.c_float, ctypes.c_float, ctypes.c_uint64,
                                  ctypes.c_int, _f32p, _f32p, _f32p, _f32p]
    L.or_run_ensemble.argtypes = [ctypes.POINTER(RunSpec), ctypes.c_size_t, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double), _f32p, ctypes.POINTER(ctypes.c_size_t)]
    L.or_run_ensemble.restype = ctypes.c_double
    L.or_time_steps.argtypes = [ctypes.c_uint32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_int, _f32p, _f32p, _f32p, _f32p, ctypes.POINTER(FlockParams),
                                ctypes.c_uint32]
    L.or_time_steps.restype = ctypes.c_double
    _lib = L
    return L


def _fp(a):
    return a.ctypes.data_as(_f32p)


def _up(a):
    return a.ctypes.data_as(_u32p)


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_up(c), _up(k), _up(o))
    return o


def step_uniforms(seed, agent_id, step):
    r1, r2 = ctypes.c_float(), ctypes.c_float()
    lib().or_step_uniforms(seed, agent_id, step, ctypes.byref(r1), ctypes.byref(r2))
    return r1.value, r2.value


def init_uniform(n, w, h, seed, heading_mode=1):
    out = np.zeros((4, max(n, 1)), dtype=np.float32)
    lib().or_init_uniform(n, w, h, seed, heading_mode, _fp(out[0]), _fp(out[1]), _fp(out[2]), _fp(out[3]))
    return out[0, :n].copy(), out[1, :n].copy(), out[2, :n].copy(), out[3, :n].copy()


def params(cohesion=1.0, avoidance=1.0, consistency=1.0, randomness=1.0, momentum=1.0,
           jump=0.7, radius=10.0, seed=0):
    return FlockParams(cohesion, avoidance, consistency, randomness, momentum, jump, radius, seed)


class Field2D:
    """Oracle Field2D with the reference method names (field_2d.rs:295-921)."""

    def __init__(self, w, h, d, toroidal):
        self._L = lib()
        self._f = self._L.or_field_new(w, h, d, int(bool(toroidal)))

    def __del__(self):
        if getattr(self, "_f", None):
            self._L.or_field_free(self._f)
            self._f = None

    def set_object_location(self, bird_id, x, y, hx=0.0, hy=0.0):
        rc = self._L.or_set_object_location(self._f, Bird(bird_id, x, y, hx, hy), x, y)
        if rc != 0:
            raise IndexError("bag index out of range")

    def set_object_locations(self, ids, x, y, hx=None, hy=None):
        """set_object_location of every bird, in array order."""
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        x, y = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, y))
        hx = None if hx is None else np.ascontiguousarray(hx, dtype=np.float32)
        hy = None if hy is None else np.ascontiguousarray(hy, dtype=np.float32)
        bad = self._L.or_set_object_locations(self._f, _up(ids), _fp(x), _fp(y),
                                              None if hx is None else _fp(hx),
                                              None if hy is None else _fp(hy), len(ids))
        if bad >= 0:
            raise IndexError(f"bag index out of range for bird {bad}")

    def neighbor_sets(self, qx, qy, dist):
        """get_neighbors_within_distance at every (qx[k], qy[k]) as CSR
        (row_ptr, ids)."""
        qx, qy = (np.ascontiguousarray(a, dtype=np.float32) for a in (qx, qy))
        n = len(qx)
        row_ptr = np.zeros(n + 1, dtype=np.uint64)
        cap = max(16, 4 * n)
        while True:
            ids = np.zeros(cap, dtype=np.uint32)
            total = self._L.or_neighbor_sets(self._f, _fp(qx), _fp(qy), n, dist,
                                             row_ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                             _up(ids), cap)
            if total <= cap:
                return row_ptr, ids[:total].copy()
            cap = int(total)

    def remove_object_location(self, bird_id, x, y):
        self._L.or_remove_object_location(self._f, bird_id, x, y)

    def lazy_update(self):
        self._L.or_lazy_update(self._f)

    def _query(self, fn, *args):
        cap = 64
        while True:
            buf = (Bird * cap)()
            n = fn(self._f, *args, buf, cap)
            if n <= cap:
                return [(b.id, b.px, b.py, b.hx, b.hy) for b in buf[:n]]
            cap = n

    def get_neighbors_within_distance(self, x, y, dist):
        return self._query(self._L.or_get_neighbors_within_distance, x, y, dist)

    def get_neighbors_within_relax_distance(self, x, y, dist):
        return self._query(self._L.or_get_neighbors_within_relax_distance, x, y, dist)

    def get_objects(self, x, y):
        return self._query(self._L.or_get_objects, x, y)

    def get_objects_unbuffered(self, x, y):
        return self._query(self._L.or_get_objects_unbuffered, x, y)

    def num_objects_at_location(self, x, y):
        return self._L.or_num_objects_at_location(self._f, x, y)

    def num_empty_bags(self):
        return self._L.or_num_empty_bags(self._f)

    def get_empty_bags(self):
        """(x, y) float32 arrays: not_discretize of every empty bag in bag
        index order (field_2d.rs:727-762)."""
        n = self._L.or_get_empty_bags(self._f, None, None, 0)
        x = np.zeros(max(n, 1), np.float32)
        y = np.zeros(max(n, 1), np.float32)
        self._L.or_get_empty_bags(self._f, _fp(x), _fp(y), n)
        return x[:n], y[:n]

    @property
    def nagents(self):
        return self._L.or_nagents(self._f)


class Flocker:
    """Oracle flockers run: agents 0..n-1 stepped by Schedule::step order."""

    def __init__(self, w, h, d, toroidal, x, y, hx, hy, p):
        self._L = lib()
        x, y, hx, hy = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, hx, hy))
        self.n = len(x)
        self.p = p
        self._fl = self._L.or_flocker_new(self.n, w, h, d, int(bool(toroidal)),
                                          _fp(x), _fp(y), _fp(hx), _fp(hy), ctypes.byref(p))

    def __del__(self):
        if getattr(self, "_fl", None):
            self._L.or_flocker_free(self._fl)
            self._fl = None

    def set_order(self, order):
        o = np.ascontiguousarray(order, dtype=np.uint32)
        assert len(o) == self.n
        self._L.or_flocker_set_order(self._fl, _up(o))

    def step(self, k=1):
        for _ in range(k):
            self._L.or_flocker_step(self._fl)

    def state(self):
        out = np.zeros((4, max(self.n, 1)), dtype=np.float32)
        self._L.or_flocker_get(self._fl, _fp(out[0]), _fp(out[1]), _fp(out[2]), _fp(out[3]))
        return out[:, :self.n].copy()


def time_steps(w, h, d, toroidal, x, y, hx, hy, p, steps):
    """Single-threaded timing of `steps` schedule steps; returns (seconds, final state)."""
    x, y, hx, hy = (np.array(a, dtype=np.float32) for a in (x, y, hx, hy))
    secs = lib().or_time_steps(len(x), w, h, d, int(bool(toroidal)), _fp(x), _fp(y), _fp(hx), _fp(hy),
                               ctypes.byref(p), steps)
    return secs, np.stack([x, y, hx, hy])


def run_ensemble(specs, nthreads, want_state=False):
    """Run independent flocker runs on a thread pool (explore_parallel!'s shape).

    specs: list of dicts with n, side, discretization, steps, params, heading_mode.
    Returns (wall_seconds, per_run_seconds, final states or None).
    """
    arr = (RunSpec * len(specs))()
    for i, s in enumerate(specs):
        arr[i] = RunSpec(s["n"], s["side"], s["discretization"], s["steps"], s["params"],
                         s.get("heading_mode", 1))
    secs = (ctypes.c_double * len(specs))()
    final = None
    offp = None
    if want_state:
        off = np.zeros(len(specs) + 1, dtype=np.uint64)
        off[1:] = np.cumsum([s["n"] for s in specs])
        final = np.zeros(int(off[-1]) * 4 + 4, dtype=np.float32)
        offc = (ctypes.c_size_t * len(specs))(*[int(v) for v in off[:-1]])
        offp = offc
    wall = lib().or_run_ensemble(arr, len(specs), nthreads, secs,
                                 _fp(final) if final is not None else None, offp)
    states = None
    if want_state:
        states = []
        for i, s in enumerate(specs):
            a = final[4 * int(off[i]): 4 * int(off[i + 1])].reshape(-1, 4).T.copy()
            states.append(a)
    return wall, list(secs), states
