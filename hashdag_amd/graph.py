"""HIP-graph replay of an edit batch and the frame traced after it (hd_frame_graph_*)."""
from __future__ import annotations

import ctypes as C

from ._lib import check, hd_edit, hd_edit_stats, hd_ray_params, lib
from .dag import HashDAG
from .tracer import DAGTracer


class FrameGraph:
    """One fused edit pass (up to max_edits editors) + paths / colours / shadows into the
    tracer's buffers, captured once and replayed per call."""

    def __init__(self, dag: HashDAG, tracer: DAGTracer | None, max_edits: int, shadow_bias: float = 1.0,
                 fog_density: float = 0.0):
        """tracer None: an edit-only graph (the pass without a trace)."""
        self.dag = dag
        self.tracer = tracer
        self.max_edits = max_edits
        h = C.c_void_p()
        if tracer is None:
            args = (0, 0, None, None)
        else:
            args = (tracer.width, tracer.height, C.c_void_p(tracer.paths.data_ptr()),
                    C.c_void_p(tracer.colors.data_ptr()))
        check(lib.hd_frame_graph_create(dag.handle, max_edits, *args, shadow_bias, fog_density, C.byref(h)),
              dag.handle)
        self._h = h

    def launch(self, edits, params: hd_ray_params | None = None):
        """Applies [(kind, (x, y, z), radius), ...] and traces `params` (no params for an
        edit-only graph); returns the edit stats and whether the captured graph ran (False: the
        eager fallback ran)."""
        arr = (hd_edit * max(len(edits), 1))()
        for i, (kind, center, radius) in enumerate(edits):
            arr[i].kind = int(kind)
            arr[i].center[:] = [float(c) for c in center]
            arr[i].radius = float(radius)
        st = hd_edit_stats()
        replayed = C.c_uint32()
        check(lib.hd_frame_graph_launch(self._h, arr, len(edits), C.byref(params) if params is not None else None,
                                        C.byref(st), C.byref(replayed)), self.dag.handle)
        return st, bool(replayed.value)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.hd_frame_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
