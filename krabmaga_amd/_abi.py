"""ctypes binding of libkrabgpu.so (include/krabgpu.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkrabgpu.so")

KG_OK = 0
KG_E_INVALID = -1
KG_E_CAPACITY = -2
KG_E_RANGE = -3
KG_E_HIP = -4
KG_E_NOMEM = -5
KG_E_STATE = -6
KG_E_ADMIT = -7

STATUS_NAMES = {
    KG_E_INVALID: "KG_E_INVALID", KG_E_CAPACITY: "KG_E_CAPACITY", KG_E_RANGE: "KG_E_RANGE",
    KG_E_HIP: "KG_E_HIP", KG_E_NOMEM: "KG_E_NOMEM", KG_E_STATE: "KG_E_STATE",
    KG_E_ADMIT: "KG_E_ADMIT",
}

# every entry point include/krabgpu.h declares
EXPORTS = [
    "kg_field_create", "kg_field_destroy", "kg_set_object_locations", "kg_init_uniform",
    "kg_remove_object_location", "kg_lazy_update", "kg_get_neighbors_within_distance",
    "kg_get_neighbors_within_relax_distance", "kg_get_objects", "kg_num_objects_at_location",
    "kg_num_empty_bags", "kg_get_empty_bags", "kg_get_random_empty_bag", "kg_iter_objects",
    "kg_nagents", "kg_count", "kg_neighbor_sets", "kg_flock_step",
    "kg_schedule_run", "kg_schedule_step_static", "kg_read_agents", "kg_field_sync", "kg_field_stream",
    "kg_field_profile", "kg_field_profile_read", "kg_field_stats",
    "kg_run_footprint", "kg_ensemble_create", "kg_ensemble_destroy", "kg_ensemble_admit",
    "kg_ensemble_retire", "kg_ensemble_run", "kg_ensemble_steps", "kg_ensemble_read",
    "kg_ensemble_sync", "kg_ensemble_stream", "kg_ensemble_stats", "kg_last_error",
]

VOID_FUNCS = ("kg_field_destroy", "kg_field_stream", "kg_last_error", "kg_ensemble_destroy",
              "kg_ensemble_stream")

KERNELS = {"flock": 0, "scan": 1, "scatter": 2, "count_append": 3, "scatter_append": 4,
           "rt_step": 5, "rt_load": 6, "sort_bins": 7}


class KgError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class FieldDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_float), ("height", ctypes.c_float),
                ("discretization", ctypes.c_float), ("toroidal", ctypes.c_int32),
                ("capacity", ctypes.c_uint32), ("bin_rows", ctypes.c_uint32),
                ("tile_cols", ctypes.c_uint32), ("tile_rows", ctypes.c_uint32)]


class FlockParams(ctypes.Structure):
    _fields_ = [("cohesion", ctypes.c_float), ("avoidance", ctypes.c_float),
                ("consistency", ctypes.c_float), ("randomness", ctypes.c_float),
                ("momentum", ctypes.c_float), ("jump", ctypes.c_float),
                ("radius", ctypes.c_float), ("reserved", ctypes.c_uint32),
                ("seed", ctypes.c_uint64)]


class RunDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_float), ("height", ctypes.c_float),
                ("discretization", ctypes.c_float), ("toroidal", ctypes.c_int32),
                ("agents", ctypes.c_uint32), ("heading_mode", ctypes.c_int32),
                ("init_seed", ctypes.c_uint64), ("params", FlockParams)]


class Tuning(ctypes.Structure):
    _fields_ = [("tile_cols", ctypes.c_uint32), ("tile_rows", ctypes.c_uint32),
                ("bin_rows", ctypes.c_uint32), ("lds_normal", ctypes.c_uint32)]


_lib = None

P = ctypes.c_void_p
# kg_object_fn(user, bag_x, bag_y, id, x, y, hx, hy)
OBJECT_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                             ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)
_szp = ctypes.POINTER(ctypes.c_size_t)


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m krabmaga_amd.build` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    I = ctypes.c_int
    L.kg_field_create.argtypes = [ctypes.POINTER(FieldDesc), ctypes.POINTER(P)]
    L.kg_field_destroy.argtypes = [P]
    L.kg_field_destroy.restype = None
    L.kg_set_object_locations.argtypes = [P, _u32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_size_t]
    L.kg_init_uniform.argtypes = [P, ctypes.c_uint32, ctypes.c_uint64, I]
    L.kg_remove_object_location.argtypes = [P, ctypes.c_uint32, ctypes.c_float, ctypes.c_float]
    L.kg_lazy_update.argtypes = [P]
    for name in ("kg_get_neighbors_within_distance", "kg_get_neighbors_within_relax_distance"):
        getattr(L, name).argtypes = [P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _u32p,
                                     ctypes.c_size_t, _szp]
    L.kg_get_objects.argtypes = [P, ctypes.c_float, ctypes.c_float, I, _u32p, ctypes.c_size_t, _szp]
    L.kg_num_objects_at_location.argtypes = [P, ctypes.c_float, ctypes.c_float, _szp]
    L.kg_num_empty_bags.argtypes = [P, _szp]
    L.kg_get_empty_bags.argtypes = [P, _f32p, _f32p, ctypes.c_size_t, _szp]
    L.kg_get_random_empty_bag.argtypes = [P, ctypes.c_uint64, _f32p, _f32p,
                                          ctypes.POINTER(ctypes.c_int)]
    L.kg_iter_objects.argtypes = [P, I, OBJECT_FN, P]
    L.kg_nagents.argtypes = [P, _szp]
    L.kg_count.argtypes = [P, I, _szp]
    L.kg_neighbor_sets.argtypes = [P, ctypes.c_float, _u32p, _u32p, ctypes.c_size_t, _szp]
    L.kg_flock_step.argtypes = [P, ctypes.POINTER(FlockParams), ctypes.c_uint64]
    L.kg_schedule_run.argtypes = [P, ctypes.POINTER(FlockParams), ctypes.c_uint64,
                                  ctypes.c_uint32, I]
    L.kg_schedule_step_static.argtypes = [ctypes.POINTER(FieldDesc), ctypes.POINTER(FlockParams),
                                          ctypes.c_uint64, P, P, P, P, P, ctypes.c_uint32, P, P,
                                          P, P, P, P]
    L.kg_read_agents.argtypes = [P, I, _u32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_size_t, _szp]
    L.kg_field_sync.argtypes = [P]
    L.kg_field_stream.argtypes = [P]
    L.kg_field_stream.restype = ctypes.c_void_p
    L.kg_field_profile.argtypes = [P, I]
    L.kg_field_profile_read.argtypes = [P, I, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_uint64)]
    L.kg_field_stats.argtypes = [P, ctypes.POINTER(ctypes.c_uint64)]
    L.kg_run_footprint.argtypes = [ctypes.POINTER(RunDesc), ctypes.POINTER(Tuning),
                                   ctypes.POINTER(ctypes.c_uint64)]
    L.kg_ensemble_create.argtypes = [ctypes.POINTER(Tuning), ctypes.POINTER(P)]
    L.kg_ensemble_destroy.argtypes = [P]
    L.kg_ensemble_destroy.restype = None
    L.kg_ensemble_admit.argtypes = [P, ctypes.POINTER(RunDesc), ctypes.c_uint32, _u32p]
    L.kg_ensemble_retire.argtypes = [P, ctypes.c_uint32]
    L.kg_ensemble_run.argtypes = [P, ctypes.c_uint32, I]
    L.kg_ensemble_steps.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]
    L.kg_ensemble_read.argtypes = [P, ctypes.c_uint32, _u32p, _f32p, _f32p, _f32p, _f32p,
                                   ctypes.c_size_t, _szp]
    L.kg_ensemble_sync.argtypes = [P]
    L.kg_ensemble_stream.argtypes = [P]
    L.kg_ensemble_stream.restype = ctypes.c_void_p
    L.kg_ensemble_stats.argtypes = [P, ctypes.POINTER(ctypes.c_uint64)]
    L.kg_last_error.argtypes = []
    L.kg_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name not in VOID_FUNCS:
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    if rc != KG_OK:
        raise KgError(rc, lib().kg_last_error().decode(errors="replace"))
    return rc


def u32p(a):
    return a.ctypes.data_as(_u32p)


def f32p(a):
    return None if a is None else a.ctypes.data_as(_f32p)
