"""ctypes bindings of liboibvh_hip.so (include/oibvh.h) and of the few HIP
runtime calls the harness needs (device buffers, events on the library's
stream).

There is no fallback: if the shared library is missing or fails to load, or
no gfx950 device is present, these calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboibvh_hip.so")

OK = 0
ERR_INVALID = 1
ERR_CAPACITY = 4
STATUS = {0: "OK", 1: "INVALID", 2: "NOT_BUILT", 3: "HIP", 4: "CAPACITY", 5: "OVERFLOW", 6: "TIMEOUT",
          7: "NO_DEVICE"}

PTR_HOST = 0
PTR_DEVICE = 1
PTR_DEVICE_BORROW = 2

DETECT_SELF = 1
DETECT_FP64_FALLBACK = 2
DETECT_ASYNC = 4
FRAME_OVERLAP = 8

# every entry point declared in include/oibvh.h: (name, restype, argtypes)
_vp, _u32, _u64, _i = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
_f32p, _u32p = C.POINTER(C.c_float), C.POINTER(C.c_uint32)


class Tuning(C.Structure):
    _fields_ = [("reduce_leaves_per_thread", C.c_uint32), ("traverse_blocks", C.c_uint32),
                ("traverse_expand", C.c_uint32), ("front_capacity", C.c_uint32), ("hits_capacity", C.c_uint32),
                ("candidate_capacity", C.c_uint32), ("traverse_schedule", C.c_uint32),
                ("traverse_local", C.c_uint32), ("traverse_local_levels", C.c_uint32),
                ("narrow_separate", C.c_uint32), ("traverse_tail", C.c_uint32),
                ("traverse_tail_steps", C.c_uint32), ("refit_blocks", C.c_uint32)]


class SceneTiming(C.Structure):
    _fields_ = [("refit_ms", C.c_double), ("detect_ms", C.c_double), ("refits", C.c_uint32),
                ("detects", C.c_uint32)]


class DetectStats(C.Structure):
    _fields_ = [("bvtt_tests", C.c_uint64), ("candidates", C.c_uint64), ("fp64_rechecks", C.c_uint64),
                ("adjacent", C.c_uint64), ("pairs", C.c_uint32), ("retries", C.c_uint32),
                ("async_errors", C.c_uint32), ("fp64_candidates", C.c_uint32)]


API = [
    ("oibvh_ctx_create", _i, [_i, _vp, C.POINTER(_vp)]),
    ("oibvh_ctx_destroy", None, [_vp]),
    ("oibvh_last_error", C.c_char_p, [_vp]),
    ("oibvh_ctx_synchronize", _i, [_vp]),
    ("oibvh_ctx_stream", _vp, [_vp]),
    ("oibvh_ctx_set_tuning", _i, [_vp, C.POINTER(Tuning)]),
    ("oibvh_ctx_get_tuning", _i, [_vp, C.POINTER(Tuning)]),
    ("oibvh_tree_create", _i, [_vp, _vp, _u32, _vp, _u32, _u32, C.POINTER(_vp)]),
    ("oibvh_tree_create_like", _i, [_vp, _vp, _u32, _u32, C.POINTER(_vp)]),
    ("oibvh_tree_destroy", None, [_vp]),
    ("oibvh_tree_build", _i, [_vp]),
    ("oibvh_tree_refit", _i, [_vp, _vp, _u32]),
    ("oibvh_tree_set_positions", _i, [_vp, _vp, _u32]),
    ("oibvh_tree_set_transform", _i, [_vp, _f32p]),
    ("oibvh_tree_set_morton_bits", _i, [_vp, _u32]),
    ("oibvh_tree_get_depth", _i, [_vp, _u32p]),
    ("oibvh_tree_get_prim_count", _i, [_vp, _u32p]),
    ("oibvh_tree_get_node_count", _i, [_vp, _u32p]),
    ("oibvh_tree_get_vertex_count", _i, [_vp, _u32p]),
    ("oibvh_tree_get_nodes", _i, [_vp, _f32p, _u64, _u32p]),
    ("oibvh_tree_get_sorted_faces", _i, [_vp, _u32p, _u64]),
    ("oibvh_tree_get_perm", _i, [_vp, _u32p, _u64]),
    ("oibvh_tree_get_positions", _i, [_vp, _f32p, _u64]),
    ("oibvh_tree_positions_device", _i, [_vp, C.POINTER(_vp)]),
    ("oibvh_scene_create", _i, [_vp, C.POINTER(_vp)]),
    ("oibvh_scene_destroy", None, [_vp]),
    ("oibvh_scene_add_tree", _i, [_vp, _vp]),
    ("oibvh_scene_detect", _i, [_vp, _u32, _u32, _u32]),
    ("oibvh_scene_get_pair_count", _i, [_vp, _u32p]),
    ("oibvh_scene_get_pairs", _i, [_vp, _u32p, _u64, _u32p]),
    ("oibvh_scene_frame", _i, [_vp, C.POINTER(_vp), _u32, _u32, _u32]),
    ("oibvh_scene_set_timing", _i, [_vp, _u32]),
    ("oibvh_scene_get_timing", _i, [_vp, C.c_void_p]),
    ("oibvh_scene_hits_device", _i, [_vp, C.POINTER(_vp), _u32p]),
    ("oibvh_scene_pairs_device", _i, [_vp, C.POINTER(_vp), _u32p]),
    ("oibvh_scene_get_pair_vertices", _i, [_vp, _vp, _u64, _u32, _u32p]),
    ("oibvh_scene_get_stats", _i, [_vp, C.POINTER(DetectStats)]),
    ("oibvh_scene_get_level_counts", _i, [_vp, _u32p, _u64, _u32p]),
    ("oibvh_scene_refit", _i, [_vp]),
    ("oibvh_scene_build", _i, [_vp]),
    ("oibvh_scene_set_transforms", _i, [_vp, _vp, _u32]),
    ("oibvh_batch_create", _i, [_vp, C.POINTER(_vp)]),
    ("oibvh_batch_destroy", None, [_vp]),
    ("oibvh_batch_add_mesh", _i, [_vp, _u32, _vp, _u32, _vp, _u32, _u32p]),
    ("oibvh_batch_add_instance", _i, [_vp, _u32, _u32, _u32p]),
    ("oibvh_batch_mesh_count", _i, [_vp, _u32p]),
    ("oibvh_batch_set_transforms", _i, [_vp, _vp, _u32]),
    ("oibvh_batch_set_morton_bits", _i, [_vp, _u32]),
    ("oibvh_batch_positions_device", _i, [_vp, _u32, C.POINTER(_vp)]),
    ("oibvh_batch_set_positions", _i, [_vp, _u32, _vp, _u32]),
    ("oibvh_batch_build", _i, [_vp]),
    ("oibvh_batch_refit", _i, [_vp]),
    ("oibvh_batch_detect", _i, [_vp, _u32, _u32, _u32]),
    ("oibvh_batch_frame", _i, [_vp, _u32, _u32, _u32]),
    ("oibvh_batch_get_pair_count", _i, [_vp, _u32p]),
    ("oibvh_batch_get_pairs", _i, [_vp, _u32p, _u64, _u32p]),
    ("oibvh_batch_get_pair_vertices", _i, [_vp, _vp, _u64, _u32, _u32p]),
    ("oibvh_batch_pairs_device", _i, [_vp, C.POINTER(_vp), _u32p]),
    ("oibvh_obj_read", _i, [_vp, C.c_char_p, _f32p, _u64, _u32p, _u64, _u32p, _u32p]),
    ("oibvh_batch_get_stats", _i, [_vp, C.POINTER(DetectStats)]),
    ("oibvh_batch_get_level_counts", _i, [_vp, _u32p, _u64, _u32p]),
    ("oibvh_batch_get_nodes", _i, [_vp, _u32, _f32p, _u64, _u32p]),
    ("oibvh_batch_get_perm", _i, [_vp, _u32, _u32p, _u64]),
    ("oibvh_batch_set_timing", _i, [_vp, _u32]),
    ("oibvh_batch_get_timing", _i, [_vp, C.c_void_p]),
    ("oibvh_build_noalloc", _i, [_vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp]),
]

NOALLOC_MAX_FACES = 1 << 20

_lib = None
_hip = None


def load():
    """Load liboibvh_hip.so (raises OSError if it is missing or broken)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, res, args in API:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class OibvhError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


def hip():
    """libamdhip64 for device buffers and events."""
    global _hip
    if _hip is None:
        H = C.CDLL("libamdhip64.so")
        H.hipMalloc.argtypes = [C.POINTER(_vp), C.c_size_t]
        H.hipFree.argtypes = [_vp]
        H.hipMemcpy.argtypes = [_vp, _vp, C.c_size_t, _i]
        H.hipMemcpyAsync.argtypes = [_vp, _vp, C.c_size_t, _i, _vp]
        H.hipStreamSynchronize.argtypes = [_vp]
        H.hipDeviceSynchronize.argtypes = []
        H.hipEventCreate.argtypes = [C.POINTER(_vp)]
        H.hipEventRecord.argtypes = [_vp, _vp]
        H.hipEventSynchronize.argtypes = [_vp]
        H.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), _vp, _vp]
        H.hipEventDestroy.argtypes = [_vp]
        H.hipGetDeviceCount.argtypes = [C.POINTER(_i)]
        H.hipSetDevice.argtypes = [_i]
        _hip = H
    return _hip


def hip_check(err: int, what: str) -> None:
    if err != 0:
        raise RuntimeError(f"{what} failed with hipError {err}")
