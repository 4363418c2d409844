"""kg_schedule_step_static, the entry point that allocates nothing, against
the same Schedule::step on a kg_field: kg_set_object_locations of every
agent, kg_lazy_update, kg_flock_step at the given schedule step,
kg_lazy_update.  Outputs are compared bit for bit in storage order.

Two calls in a row on one GPU share the static scratch array; the second
call has a smaller input, so stale bin counts or offsets of the first would
show up in its result.
"""
import ctypes

import numpy as np
import pytest

D = 10.0 / 1.5


def _lib():
    from krabmaga_amd import _abi
    return _abi.lib(), _abi


def agents(n, side, seed):
    rng = np.random.default_rng(seed)
    ids = rng.permutation(np.arange(n, dtype=np.uint32) * 3 + 7).astype(np.uint32)
    x = (rng.random(n) * side).astype(np.float32)
    y = (rng.random(n) * side).astype(np.float32)
    a = rng.random(n) * 2 * np.pi
    return {"id": ids, "x": x, "y": y, "hx": np.cos(a).astype(np.float32),
            "hy": np.sin(a).astype(np.float32)}


def params(seed):
    from krabmaga_amd import FlockParams
    return FlockParams(cohesion=1.0, avoidance=1.0, consistency=1.0, randomness=1.0,
                       momentum=1.0, jump=0.7, radius=10.0, seed=seed)


class DeviceArrays:
    """Device copies of agent arrays, allocated by the caller (this test)
    through the HIP runtime the library itself uses."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def upload(self, a):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


KEYS = ("id", "x", "y", "hx", "hy")


def via_static(desc, p, step, a):
    L, _abi = _lib()
    n = len(a["id"])
    mem = DeviceArrays()
    try:
        src = {k: mem.upload(a[k]) for k in KEYS}
        dst = {k: mem.upload(np.zeros_like(a[k])) for k in KEYS}
        rc = L.kg_schedule_step_static(
            ctypes.byref(desc), ctypes.byref(p.c()), step,
            src["id"], src["x"], src["y"], src["hx"], src["hy"], n,
            dst["id"], dst["x"], dst["y"], dst["hx"], dst["hy"], None)
        _abi.check(rc)
        return {k: mem.download(dst[k], a[k]) for k in KEYS}
    finally:
        mem.free()


def via_field(desc, p, step, a):
    from krabmaga_amd import Field2D
    f = Field2D(desc.width, desc.height, desc.discretization, bool(desc.toroidal), desc.capacity)
    f.set_object_locations(a["id"], a["x"], a["y"], a["hx"], a["hy"])
    f.lazy_update()
    f.flock_step(p, step)
    f.lazy_update()
    out = f.read_agents(0)
    f.close()
    return out


def same_bits(u, v):
    for k in ("id", "x", "y", "hx", "hy"):
        assert u[k].shape == v[k].shape, k
        assert np.array_equal(u[k].view(np.uint32), v[k].view(np.uint32)), k


@pytest.mark.gpu
def test_static_step_matches_the_field_step_twice_smaller_second():
    _, _abi = _lib()
    cases = [
        (20000, 1000.0, True, 5, 17),   # 0.02 birds per unit^2
        (1000, 200.0, True, 0, 0),      # the 1,000-boid configuration, step 0
        (3000, 300.0, False, 2, 9),     # not toroidal
    ]
    for n, side, tor, step, seed in cases:
        print(f"static step: n={n} side={side} toroidal={tor} step={step}", flush=True)
        desc = _abi.FieldDesc(side, side, D, int(tor), n, 0, 0, 0)
        a = agents(n, side, seed)
        p = params(seed)
        got = via_static(desc, p, step, a)
        want = via_field(desc, p, step, a)
        same_bits(got, want)
        assert sorted(got["id"].tolist()) == sorted(a["id"].tolist())


def test_static_step_rejects_sizes_beyond_the_scratch():
    L, _abi = _lib()
    # 4M agents need about 160 MB of scratch for the binned copies alone
    n = 4 << 20
    desc = _abi.FieldDesc(40960.0, 40960.0, D, 1, n, 0, 0, 0)
    p = params(0)
    rc = L.kg_schedule_step_static(ctypes.byref(desc), ctypes.byref(p.c()), 0, None, None, None,
                                   None, None, n, None, None, None, None, None, None)
    assert rc == _abi.KG_E_CAPACITY
    assert b"scratch" in L.kg_last_error()
    desc = _abi.FieldDesc(200.0, 200.0, D, 1, 100, 0, 0, 0)
    rc = L.kg_schedule_step_static(ctypes.byref(desc), ctypes.byref(p.c()), 0, None, None, None,
                                   None, None, 101, None, None, None, None, None, None)
    assert rc == _abi.KG_E_CAPACITY
