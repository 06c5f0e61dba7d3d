"""Replays: the reference's CSV action streams (ReplayManager, replay.h:12-635; replay.cpp:1-101),
parsed by the library (hd_replay_parse) and turned into a served request stream.

Frame k of a replay is every action up to its k-th EndFrame (ReplayManager::replay_frame,
replay.h:497-513): camera moves, edits (scaled from the replay's depth to the scene's,
replay.cpp:5), undo / redo. Served, a frame becomes its edits as one edit request, each undo /
redo as its own request, and the frame's trace at its end, all arriving at k / fps.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import (HD_REPLAY_EDIT, HD_REPLAY_END_FRAME, HD_REPLAY_REDO, HD_REPLAY_SET_LOCATION, HD_REPLAY_SET_ROTATION,
                   HD_REPLAY_UNDO, HD_REQ_EDIT, HD_REQ_REDO, HD_REQ_TRACE, HD_REQ_UNDO, check, hd_replay_action, lib)
from .tracer import CameraView, DAGInfo, ray_params

COPY = 5
REPLAY_DEPTH = 17  # the depth the reference's replays were recorded at (typedefs.h:24-29)


@dataclass
class ReplayAction:
    type: int
    frame: int
    location: tuple | None = None
    rotation: tuple | None = None
    edit: tuple | None = None   # (kind, center, radius[, source, dest, transform, swirl_period])


def _edit_tuple(e) -> tuple:
    center = tuple(float(v) for v in e.center)
    if e.kind != COPY:
        return (int(e.kind), center, float(e.radius))
    T = np.array(list(e.transform), np.float64).reshape(3, 3)
    return (COPY, center, float(e.radius), tuple(float(v) for v in e.source), tuple(float(v) for v in e.dest), T,
            float(e.swirl_period) if e.swirl else None)


class Replay:
    """A parsed replay: `actions` (ReplayAction) and the number of `frames`."""

    def __init__(self, actions: list[ReplayAction], frames: int):
        self.actions = actions
        self.frames = frames

    @staticmethod
    def parse(text, scene_depth: int, replay_depth: int = REPLAY_DEPTH) -> "Replay":
        data = text.encode() if isinstance(text, str) else bytes(text)
        h = C.c_void_p()
        check(lib.hd_replay_parse(data, len(data), scene_depth, replay_depth, C.byref(h)))
        try:
            arr = C.POINTER(hd_replay_action)()
            n, frames = C.c_uint64(), C.c_uint32()
            check(lib.hd_replay_actions(h, C.byref(arr), C.byref(n), C.byref(frames)))
            out = []
            for i in range(n.value):
                a = arr[i]
                r = ReplayAction(int(a.type), int(a.frame))
                if a.type == HD_REPLAY_SET_LOCATION:
                    r.location = tuple(a.location)
                elif a.type == HD_REPLAY_SET_ROTATION:
                    r.rotation = tuple(a.rotation)
                elif a.type == HD_REPLAY_EDIT:
                    r.edit = _edit_tuple(a.edit)
                out.append(r)
            return Replay(out, int(frames.value))
        finally:
            lib.hd_replay_free(h)

    @staticmethod
    def load(path: str, scene_depth: int, replay_depth: int = REPLAY_DEPTH) -> "Replay":
        with open(path, "rb") as f:
            return Replay.parse(f.read(), scene_depth, replay_depth)

    def counts(self) -> dict:
        names = {HD_REPLAY_EDIT: "edits", HD_REPLAY_UNDO: "undo", HD_REPLAY_REDO: "redo", HD_REPLAY_END_FRAME: "frames"}
        out = {v: 0 for v in names.values()}
        for a in self.actions:
            if a.type in names:
                out[names[a.type]] += 1
        return out

    def cameras(self) -> list[CameraView]:
        """The camera at the end of every frame (the last SetLocation / SetRotation so far)."""
        pos, rot = (0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
        out = []
        for a in self.actions:
            if a.type == HD_REPLAY_SET_LOCATION:
                pos = a.location
            elif a.type == HD_REPLAY_SET_ROTATION:
                rot = a.rotation
            elif a.type == HD_REPLAY_END_FRAME:
                out.append(CameraView(pos, rot))
        return out


def replay_requests(replay: Replay, width: int, height: int, levels: int, fps: float = 60.0, params=None,
                    info: DAGInfo | None = None):
    """The replay as a served request stream (serve.Request list). Trace requests use `params`
    (hd_ray_params, cycled per frame) when given, else the replay's own camera of the frame with
    scene bounds `info`."""
    from .serve import Request

    if params is None:
        info = info or DAGInfo.unit(levels)
        cams = replay.cameras()
    reqs, pending = [], []
    for a in replay.actions:
        t = a.frame / fps
        if a.type == HD_REPLAY_EDIT:
            pending.append(a.edit)
        elif a.type in (HD_REPLAY_UNDO, HD_REPLAY_REDO):
            if pending:
                reqs.append(Request(t, HD_REQ_EDIT, edits=pending))
                pending = []
            reqs.append(Request(t, HD_REQ_UNDO if a.type == HD_REPLAY_UNDO else HD_REQ_REDO))
        elif a.type == HD_REPLAY_END_FRAME:
            if pending:
                reqs.append(Request(t, HD_REQ_EDIT, edits=pending))
                pending = []
            prm = params[a.frame % len(params)] if params is not None else \
                ray_params(cams[a.frame], info, levels, width, height)
            reqs.append(Request(t, HD_REQ_TRACE, width, height, prm))
    if pending:
        reqs.append(Request(replay.frames / fps, HD_REQ_EDIT, edits=pending))
    return reqs
