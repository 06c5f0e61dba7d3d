"""Camera setup and the DAGTracer surface (dag_tracer.h:9-42): resolve_paths,
resolve_colors, resolve_shadows over a device-resident HashDAG. Images are device buffers
(torch tensors on the GPU): paths as W*H uint4 (int32 view), colours as W*H RGBA8 words.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, hd_ray_params, lib, ptr
from .dag import HashDAG


@dataclass
class CameraView:
    """CameraView (camera_view.h:8-27): FP64 position and rotation rows (right = -row0,
    up = row1, forward = row2); vertical fov 60 degrees."""
    position: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)

    @staticmethod
    def look(position, yaw: float, pitch: float) -> "CameraView":
        f = np.array([math.cos(pitch) * math.sin(yaw), math.sin(pitch), math.cos(pitch) * math.cos(yaw)])
        world_up = np.array([0.0, 1.0, 0.0])
        right = np.cross(f, world_up)
        right /= np.linalg.norm(right)
        up = np.cross(right, f)
        rot = (*(-right), *up, *f)
        return CameraView(tuple(float(v) for v in position), tuple(float(v) for v in rot))


@dataclass
class DAGInfo:
    """DAGInfo (dag_info.h:5-9): scene bounds."""
    bounds_min: tuple = (0.0, 0.0, 0.0)
    bounds_max: tuple = (1.0, 1.0, 1.0)

    @staticmethod
    def unit(levels: int) -> "DAGInfo":
        e = float(1 << levels)
        return DAGInfo((0.0, 0.0, 0.0), (e, e, e))


def ray_params(camera: CameraView, info: DAGInfo, levels: int, width: int, height: int) -> hd_ray_params:
    """get_trace_params (dag_tracer.cu:71-113), computed by the library in FP64."""
    pos = (C.c_double * 3)(*camera.position)
    rot = (C.c_double * 9)(*camera.rotation)
    bmin = (C.c_double * 3)(*info.bounds_min)
    bmax = (C.c_double * 3)(*info.bounds_max)
    out = hd_ray_params()
    check(lib.hd_ray_params_from_camera(pos, rot, bmin, bmax, levels, width, height, C.byref(out)))
    return out


def _stream_ptr():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class DAGTracer:
    """Owns the paths and colours images for one resolution (dag_tracer.cu:9-45)."""

    def __init__(self, width: int = 1920, height: int = 1080, device: str = "cuda:0"):
        import torch
        self.width, self.height = width, height
        self.paths = torch.zeros((height, width, 4), dtype=torch.int32, device=device)
        self.colors = torch.zeros((height, width), dtype=torch.int32, device=device)

    def resolve_paths(self, camera: CameraView, info: DAGInfo, dag: HashDAG, params: hd_ray_params | None = None):
        p = params or ray_params(camera, info, dag.geometry.levels, self.width, self.height)
        check(lib.hd_trace_paths(dag.handle, C.byref(p), self.width, self.height, C.c_void_p(self.paths.data_ptr()),
                                 _stream_ptr()), dag.handle)
        return p

    def resolve_colors(self, dag: HashDAG):
        check(lib.hd_trace_colors(dag.handle, self.width, self.height, C.c_void_p(self.paths.data_ptr()),
                                  C.c_void_p(self.colors.data_ptr()), _stream_ptr()), dag.handle)

    def resolve_shadows(self, camera: CameraView, info: DAGInfo, dag: HashDAG, shadow_bias: float = 1.0,
                        fog_density: float = 0.0, params: hd_ray_params | None = None):
        p = params or ray_params(camera, info, dag.geometry.levels, self.width, self.height)
        check(lib.hd_trace_shadows(dag.handle, C.byref(p), self.width, self.height, C.c_void_p(self.paths.data_ptr()),
                                   C.c_void_p(self.colors.data_ptr()), shadow_bias, fog_density, _stream_ptr()),
              dag.handle)

    def count_bytes(self, dag: HashDAG, params: hd_ray_params, shadow_bias: float = 1.0, fog_density: float = 0.0):
        """Algorithmic bytes of one frame per kernel (paths, colours, shadows)."""
        out = np.zeros(3, np.uint64)
        check(lib.hd_trace_count_bytes(dag.handle, C.byref(params), self.width, self.height,
                                       C.c_void_p(self.paths.data_ptr()), C.c_void_p(self.colors.data_ptr()),
                                       shadow_bias, fog_density, ptr(out, C.c_uint64)), dag.handle)
        return [int(v) for v in out]

    def set_tile(self, dag: HashDAG, tile_w: int, tile_h: int) -> None:
        check(lib.hd_set_trace_tile(dag.handle, tile_w, tile_h), dag.handle)

    def set_fetch(self, dag: HashDAG, fetch: int = 0) -> None:
        """Node fetch strategy: 0 = child pointer then header, 1 = whole node."""
        check(lib.hd_set_trace_fetch(dag.handle, fetch), dag.handle)


    def set_schedule(self, dag: HashDAG, schedule, param: int = 16, resume_to: int = 0) -> None:
        """0 (or False): one wavefront per pixel tile; 1 (or True): persistent wavefronts fed
        from tile queues, refilling idle lanes once `param` of them are idle; 2: one wavefront
        per tile, walks beyond `param` steps searched a wavefront per ray; 3: one wavefront per
        tile to `param` steps, the unfinished walks continued by compacted wavefronts to
        `resume_to` steps, the rest searched a wavefront per ray (or `param` already packed as
        A | B << 16 when `resume_to` is 0)."""
        if int(schedule) == 3 and resume_to:
            param = (param & 0xFFFF) | (resume_to << 16)
        check(lib.hd_set_trace_schedule(dag.handle, int(schedule), param), dag.handle)


def launch_config(dag: HashDAG) -> dict:
    out = (C.c_uint32 * 3)()
    check(lib.hd_get_trace_launch(dag.handle, out), dag.handle)
    sch = np.zeros(2, np.uint32)
    check(lib.hd_get_trace_schedule(dag.handle, ptr(sch)), dag.handle)
    c1, c2 = int(sch[1]) & 0xFFFF, int(sch[1]) >> 16
    schedule = {0: "per tile", 1: f"persistent (refill at {int(sch[1])})",
                2: f"per tile, walks beyond {int(sch[1])} steps a wavefront per ray",
                3: f"per tile to {c1} steps, resumed by compacted wavefronts to {c2}, the rest a wavefront per "
                   f"ray"}[int(sch[0])]
    key = {0: "tile", 1: f"persist:{int(sch[1])}", 2: f"beam:{int(sch[1])}", 3: f"resume:{c1}:{c2}"}[int(sch[0])]
    return {"tile": f"{out[0]}x{out[1]}", "fetch": ["pointer", "node"][out[2]], "schedule": schedule,
            "schedule_key": key}


def autotune(tracer: "DAGTracer", dag: HashDAG, params_list, reps: int = 2) -> dict:
    """Picks the tracer's wavefront tile shape and node fetch strategy by timing full frames
    over the given cameras (hd_autotune_trace); the choice stays in the DAG context for later
    traces."""
    arr = (hd_ray_params * len(params_list))(*params_list)
    bw, bh, ms = C.c_uint32(), C.c_uint32(), C.c_float()
    check(lib.hd_autotune_trace(dag.handle, arr, len(params_list), tracer.width, tracer.height,
                                C.c_void_p(tracer.paths.data_ptr()), C.c_void_p(tracer.colors.data_ptr()), reps,
                                C.byref(bw), C.byref(bh), C.byref(ms)), dag.handle)
    return {**launch_config(dag), "ms_per_frame": round(ms.value, 4)}
