"""ctypes binding of libhashdag_hip.so (the C ABI declared in include/hashdag_hip.h).

The library is loaded from this package directory. There is no fallback: if the HIP
library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhashdag_hip.so")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
u8p = C.POINTER(C.c_uint8)
dblp = C.POINTER(C.c_double)


class HashDAGError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


HD_OK = 0
HD_ERR_INVALID = -1
HD_ERR_HIP = -2
HD_ERR_BUCKET_FULL = -3
HD_ERR_POOL_FULL = -4
HD_ERR_OOM = -5
HD_ERR_COLOR_FULL = -6
HD_ERR_HEADER_OVERFLOW = -7


class hd_config(C.Structure):
    _fields_ = [
        ("levels", C.c_uint32),
        ("page_size", C.c_uint32),
        ("num_top_levels", C.c_uint32),
        ("bucket_bits_top", C.c_uint32),
        ("bucket_bits_low", C.c_uint32),
        ("bucket_size_top", C.c_uint32),
        ("bucket_size_low", C.c_uint32),
        ("color_tree_levels", C.c_uint32),
        ("pool_pages", C.c_uint32),
        ("device", C.c_uint32),
        ("color_arena_words", C.c_uint64),
        ("color_nodes_cap", C.c_uint32),
        ("color_leaves_cap", C.c_uint32),
    ]


class hd_ray_params(C.Structure):
    _fields_ = [
        ("camera_position", C.c_double * 3),
        ("ray_min", C.c_double * 3),
        ("ray_ddx", C.c_double * 3),
        ("ray_ddy", C.c_double * 3),
    ]

    def as_list(self) -> list[float]:
        return [*self.camera_position, *self.ray_min, *self.ray_ddx, *self.ray_ddy]


class hd_image_info(C.Structure):
    _fields_ = [
        ("root", C.c_uint32),
        ("pool_top", C.c_uint32),
        ("pool_pages", C.c_uint32),
        ("levels", C.c_uint32),
        ("total_pages", C.c_uint64),
        ("total_buckets", C.c_uint64),
        ("total_virtual", C.c_uint64),
    ]


class hd_terrain_params(C.Structure):
    _fields_ = [
        ("levels", C.c_uint32),
        ("footprint", C.c_uint32),
        ("seed", C.c_uint64),
        ("octaves", C.c_uint32),
        ("amplitude", C.c_double),
        ("frequency", C.c_double),
        ("color_tree_levels", C.c_uint32),
        ("colors", C.c_uint32),
        ("color_seed", C.c_uint32),
        ("threads", C.c_uint32),
    ]


class hd_edit(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("center", C.c_float * 3), ("radius", C.c_float), ("source", C.c_float * 3),
                ("dest", C.c_float * 3), ("swirl", C.c_uint32), ("swirl_period", C.c_float),
                ("transform", C.c_double * 9)]


class hd_replay_action(C.Structure):
    _fields_ = [("type", C.c_uint32), ("frame", C.c_uint32), ("location", C.c_double * 3),
                ("rotation", C.c_double * 9), ("edit", hd_edit)]


class hd_edit_stats(C.Structure):
    _fields_ = [("num_voxels", C.c_uint64), ("new_words", C.c_uint64), ("leaves_rebuilt", C.c_uint64),
                ("root", C.c_uint32), ("passes", C.c_uint32), ("colors_encoded", C.c_uint64)]


HD_REQ_TRACE, HD_REQ_EDIT, HD_REQ_UNDO, HD_REQ_REDO = 0, 1, 2, 3

(HD_REPLAY_SET_LOCATION, HD_REPLAY_SET_ROTATION, HD_REPLAY_SET_TOOL, HD_REPLAY_EDIT, HD_REPLAY_UNDO, HD_REPLAY_REDO,
 HD_REPLAY_END_FRAME) = range(7)


class hd_request(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("kind", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("first_edit", C.c_uint32), ("n_edits", C.c_uint32), ("pad", C.c_uint32),
                ("params", hd_ray_params)]


class hd_request_result(C.Structure):
    _fields_ = [("start_s", C.c_double), ("done_s", C.c_double), ("checksum", C.c_uint64), ("voxels", C.c_uint64),
                ("batch", C.c_uint32), ("status", C.c_int32), ("gc", C.c_uint32), ("pad", C.c_uint32)]


class hd_mem_info_t(C.Structure):
    _fields_ = [("pool_pages", C.c_uint64), ("pool_top", C.c_uint64), ("reusable_pages", C.c_uint64),
                ("free_pages", C.c_uint64), ("used_words", C.c_uint64), ("arena_words", C.c_uint64),
                ("arena_top", C.c_uint64)]


class hd_serve_config(C.Structure):
    _fields_ = [("trace_slots", C.c_uint32), ("max_batch_edits", C.c_uint32), ("realtime", C.c_uint32),
                ("keep_undo", C.c_uint32), ("reserve_pages", C.c_uint64), ("shadow_bias", C.c_float),
                ("fog_density", C.c_float)]


_SIGS = {
    "hd_serve_run": (C.c_int32, [C.c_void_p, C.POINTER(hd_request), C.c_uint32, C.POINTER(hd_edit), C.c_uint32,
                                 C.POINTER(hd_serve_config), C.POINTER(hd_request_result)]),
    "hd_frame_graph_create": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                          C.c_float, C.c_float, C.POINTER(C.c_void_p)]),
    "hd_frame_graph_launch": (C.c_int32, [C.c_void_p, C.POINTER(hd_edit), C.c_uint32, C.POINTER(hd_ray_params),
                                          C.POINTER(hd_edit_stats), u32p]),
    "hd_frame_graph_destroy": (C.c_int32, [C.c_void_p]),
    "hd_mem_info": (C.c_int32, [C.c_void_p, C.POINTER(hd_mem_info_t)]),
    "hd_frame_checksum": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, u64p, C.c_void_p]),
    "hd_edit_batch": (C.c_int32, [C.c_void_p, C.POINTER(hd_edit), C.c_uint32, C.c_uint32, C.POINTER(hd_edit_stats),
                                  C.c_void_p]),
    "hd_edit_batch_static": (C.c_int32, [C.c_void_p, C.POINTER(hd_edit), C.c_uint32, C.POINTER(hd_edit_stats),
                                         C.c_void_p]),
    "hd_edit_count_bytes": (C.c_int32, [C.c_void_p, C.c_uint32, u64p]),
    "hd_set_edit_launch": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]),
    "hd_get_edit_launch": (C.c_int32, [C.c_void_p, u32p, u32p, u32p]),
    "hd_autotune_edit_launch": (C.c_int32, [C.c_void_p, C.POINTER(hd_edit), C.c_uint32, C.c_uint32, u32p,
                                            C.POINTER(C.c_float)]),
    "hd_gc": (C.c_int32, [C.c_void_p, C.c_uint32, u64p, C.c_void_p]),
    "hd_write_basic_dag_file": (C.c_int32, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint32, u32p,
                                            C.c_uint64]),
    "hd_read_basic_dag_file": (C.c_int32, [C.c_char_p, C.POINTER(C.c_double), u32p, u32p, C.c_uint64, u64p]),
    "hd_write_compressed_colors_file": (C.c_int32, [C.c_char_p, C.c_uint32, u64p, C.c_uint64, u32p, C.c_uint64, u64p,
                                                    C.c_uint64, u64p, C.c_uint64]),
    "hd_read_compressed_colors_file": (C.c_int32, [C.c_char_p, u32p, u64p, u64p, u32p, u64p, u64p]),
    "hd_basic_dag_counts": (C.c_int32, [u32p, C.c_uint64, C.c_uint32, C.c_uint32, u32p, u64p, u64p]),
    "hd_colors_from_basic_dag": (C.c_int32, [u32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, u64p, C.c_uint64,
                                             u32p, u32p, u64p, u64p, u64p]),
    "hd_save_hash_dag_file": (C.c_int32, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hd_load_hash_dag_file": (C.c_int32, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    "hd_checkpoint_save": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "hd_checkpoint_load": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "hd_validate_image": (C.c_int32, [C.POINTER(hd_config), u32p, u32p, u32p, u32p, C.c_uint32, C.c_uint32,
                                      C.c_char_p, C.c_uint32]),
    "hd_replay_parse": (C.c_int32, [C.c_char_p, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "hd_replay_actions": (C.c_int32, [C.c_void_p, C.POINTER(C.POINTER(hd_replay_action)), u64p,
                                      C.POINTER(C.c_uint32)]),
    "hd_replay_free": (None, [C.c_void_p]),
    "hd_gc_roots": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32, u64p,
                                C.c_void_p]),
    "hd_undo": (C.c_int32, [C.c_void_p]),
    "hd_redo": (C.c_int32, [C.c_void_p]),
    "hd_undo_clear": (C.c_int32, [C.c_void_p]),
    "hd_undo_depth": (C.c_int32, [C.c_void_p, u32p, u32p]),
    "hd_reserve_edit_scratch": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64]),
    "hd_default_config": (None, [C.c_uint32, C.POINTER(hd_config)]),
    "hd_create": (C.c_int32, [C.POINTER(hd_config), C.POINTER(C.c_void_p)]),
    "hd_destroy": (C.c_int32, [C.c_void_p]),
    "hd_last_error": (C.c_char_p, [C.c_void_p]),
    "hd_load_basic_dag": (C.c_int32, [C.c_void_p, u32p, C.c_uint64]),
    "hd_load_colors": (C.c_int32, [C.c_void_p, u32p, C.c_uint64, u64p, C.c_uint64, u64p, C.c_uint64, u32p,
                                   C.c_uint64, u64p, C.c_uint64]),
    "hd_image_info_get": (C.c_int32, [C.c_void_p, C.POINTER(hd_image_info)]),
    "hd_download_image": (C.c_int32, [C.c_void_p, u32p, u32p, u32p, u32p]),
    "hd_upload_image": (C.c_int32, [C.c_void_p, u32p, u32p, u32p, u32p, C.c_uint32, C.c_uint32]),
    "hd_get_root": (C.c_int32, [C.c_void_p, u32p]),
    "hd_set_root": (C.c_int32, [C.c_void_p, C.c_uint32]),
    "hd_colors_info": (C.c_int32, [C.c_void_p, u64p, u64p, u64p]),
    "hd_download_color_nodes": (C.c_int32, [C.c_void_p, u32p]),
    "hd_color_leaf_sizes": (C.c_int32, [C.c_void_p, C.c_uint32, u64p]),
    "hd_download_color_leaf": (C.c_int32, [C.c_void_p, C.c_uint32, u64p, u32p, u64p]),
    "hd_ray_params_from_camera": (C.c_int32, [dblp, dblp, dblp, dblp, C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.POINTER(hd_ray_params)]),
    "hd_trace_paths": (C.c_int32, [C.c_void_p, C.POINTER(hd_ray_params), C.c_uint32, C.c_uint32, C.c_void_p,
                                   C.c_void_p]),
    "hd_trace_colors": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hd_trace_shadows": (C.c_int32, [C.c_void_p, C.POINTER(hd_ray_params), C.c_uint32, C.c_uint32, C.c_void_p,
                                     C.c_void_p, C.c_float, C.c_float, C.c_void_p]),
    "hd_autotune_trace": (C.c_int32, [C.c_void_p, C.POINTER(hd_ray_params), C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_void_p, C.c_void_p, C.c_uint32, u32p, u32p, C.POINTER(C.c_float)]),
    "hd_set_trace_tile": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "hd_set_trace_fetch": (C.c_int32, [C.c_void_p, C.c_uint32]),
    "hd_set_trace_schedule": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "hd_get_trace_schedule": (C.c_int32, [C.c_void_p, u32p]),
    "hd_get_trace_launch": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "hd_trace_count_bytes": (C.c_int32, [C.c_void_p, C.POINTER(hd_ray_params), C.c_uint32, C.c_uint32, C.c_void_p,
                                         C.c_void_p, C.c_float, C.c_float, u64p]),
    "hd_stream_copy": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "hd_stream_read": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
    "hd_scene_generate": (C.c_int32, [C.POINTER(hd_terrain_params), C.POINTER(C.c_void_p)]),
    "hd_scene_free": (None, [C.c_void_p]),
    "hd_scene_sizes": (C.c_int32, [C.c_void_p, u64p]),
    "hd_scene_copy": (C.c_int32, [C.c_void_p, u32p, u32p, u64p, u64p, u32p, u64p]),
    "hd_scene_level_nodes": (C.c_int32, [C.c_void_p, u64p]),
    "hd_terrain_height": (C.c_int32, [C.POINTER(hd_terrain_params), C.c_uint32, C.c_uint32]),
}

EXPORTED = sorted(_SIGS)


def _load() -> C.CDLL:
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 / libhsa-runtime64.so.1;
    # loading torch first makes the library's dependencies resolve to those (same sonames)
    # instead of pulling a second runtime from /opt/rocm, which cannot share the device.
    import torch  # noqa: F401

    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP library has no fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(status: int, ctx=None) -> None:
    if status != HD_OK:
        msg = lib.hd_last_error(ctx)
        raise HashDAGError(status, msg.decode() if msg else "error")


def ptr(arr, ctype=C.c_uint32):
    """ctypes pointer to a contiguous numpy array (None passes NULL)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))
