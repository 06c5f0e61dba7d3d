"""Request serving on one device-resident HashDAG (continuous batching in hd_serve_run).

A request stream mixes trace requests (a camera and a frame size) and edit requests (a run of
sphere / box editors). Requests take effect in arrival order, so every trace result is the
same whether the request is served alone or in a mix; the scheduler fuses runs of edit
requests into one edit pass and runs runs of trace requests concurrently.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import (HD_REQ_EDIT, HD_REQ_REDO, HD_REQ_TRACE, HD_REQ_UNDO, check, hd_edit, hd_ray_params, hd_request,
                   hd_request_result, hd_serve_config, lib)
from .dag import HashDAG, fill_edit
from .scene import TerrainParams, terrain_height
from .tracer import DAGInfo, ray_params
from .workload import BOX_ADD, BOX_CARVE, SPHERE_ADD, SPHERE_CARVE, cameras

RESOLUTIONS = ((640, 360), (1280, 720), (1920, 1080), (3840, 2160))


@dataclass
class Request:
    arrival_s: float
    kind: int                     # HD_REQ_TRACE / HD_REQ_EDIT / HD_REQ_UNDO / HD_REQ_REDO
    width: int = 0
    height: int = 0
    params: hd_ray_params | None = None
    edits: list = field(default_factory=list)  # [(kind, (x, y, z), radius), ...]


def request_stream(params: TerrainParams, n: int, rate: float, seed: int = 5, trace_fraction: float = 0.6,
                   resolutions=RESOLUTIONS, edits_per_request=(1, 64), rmin: float = 2.0,
                   rmax: float = 128.0) -> list[Request]:
    """Poisson arrivals at `rate` requests/s: trace requests (uniform resolution, seeded camera)
    with probability trace_fraction, else edit requests of 1-64 sphere / box editors (add or
    carve) with log-uniform radius, centred at seeded integer surface points."""
    rng = np.random.default_rng(seed)
    F = params.footprint or (1 << params.levels)
    cams = cameras(params, 64, seed=seed + 1)
    info = DAGInfo.unit(params.levels)
    out, t = [], 0.0
    for _ in range(n):
        t += float(rng.exponential(1.0 / rate))
        if rng.random() < trace_fraction:
            W, H = resolutions[int(rng.integers(0, len(resolutions)))]
            cam = cams[int(rng.integers(0, len(cams)))]
            out.append(Request(t, HD_REQ_TRACE, W, H, ray_params(cam, info, params.levels, W, H)))
        else:
            k = int(rng.integers(edits_per_request[0], edits_per_request[1] + 1))
            eds = []
            for _ in range(k):
                x, z = (int(v) for v in rng.integers(0, F, size=2))
                y = terrain_height(params, x, z)
                r = float(np.float32(math.exp(rng.uniform(math.log(rmin), math.log(rmax)))))
                kind = [SPHERE_ADD, SPHERE_CARVE, BOX_ADD, BOX_CARVE][int(rng.integers(0, 4))]
                eds.append((kind, (float(x), float(y), float(z)), r))
            out.append(Request(t, HD_REQ_EDIT, edits=eds))
    return out


def serve(dag: HashDAG, requests: list[Request], trace_slots: int = 4, max_batch_edits: int = 1 << 16,
          realtime: bool = True, keep_undo: bool = False, reserve_pages: int = 1024, shadow_bias: float = 1.0,
          fog_density: float = 0.0) -> np.ndarray:
    """Runs the stream; returns one record per request: start_s, done_s, checksum, voxels,
    batch, status (see hd_request_result)."""
    n = len(requests)
    reqs = (hd_request * max(n, 1))()
    flat = []
    for i, r in enumerate(requests):
        q = reqs[i]
        q.arrival_s = r.arrival_s
        q.kind = r.kind
        if r.kind == HD_REQ_TRACE:
            q.width, q.height = r.width, r.height
            q.params = r.params
        elif r.kind == HD_REQ_EDIT:
            q.first_edit = len(flat)
            q.n_edits = len(r.edits)
            flat.extend(r.edits)
    edits = (hd_edit * max(len(flat), 1))()
    for i, ed in enumerate(flat):
        fill_edit(edits[i], ed)
    cfg = hd_serve_config(trace_slots, max_batch_edits, 1 if realtime else 0, 1 if keep_undo else 0, reserve_pages,
                          shadow_bias, fog_density)
    res = (hd_request_result * max(n, 1))()
    check(lib.hd_serve_run(dag.handle, reqs, n, edits, len(flat), C.byref(cfg), res), dag.handle)
    dt = np.dtype([("start_s", "f8"), ("done_s", "f8"), ("checksum", "u8"), ("voxels", "u8"), ("batch", "u4"),
                   ("status", "i4"), ("gc", "u4"), ("pad", "u4")])
    return np.frombuffer(bytes(res), dtype=dt, count=n).copy()


def frame_checksum(dag: HashDAG, colors, width: int, height: int, stream=None) -> int:
    out = C.c_uint64()
    check(lib.hd_frame_checksum(dag.handle, C.c_void_p(colors.data_ptr()), width, height, C.byref(out), stream),
          dag.handle)
    return out.value


def latency_summary(requests: list[Request], results: np.ndarray) -> dict:
    """p50 / p99 latency (completion - arrival) per request kind, and throughput."""
    arrival = np.array([r.arrival_s for r in requests])
    lat = results["done_s"] - arrival
    kinds = np.array([r.kind for r in requests])
    span = float(results["done_s"].max() - min(arrival.min(), results["start_s"].min()))
    out = {"requests": len(requests), "span_s": round(span, 4), "requests_per_s": round(len(requests) / span, 2),
           "batches": int(results["batch"].max()) + 1 if len(results) else 0}
    for name, k in (("trace", HD_REQ_TRACE), ("edit", HD_REQ_EDIT), ("undo", HD_REQ_UNDO), ("redo", HD_REQ_REDO)):
        sel = lat[kinds == k]
        if sel.size:
            out[f"{name}_p50_ms"] = round(float(np.percentile(sel, 50)) * 1e3, 3)
            out[f"{name}_p99_ms"] = round(float(np.percentile(sel, 99)) * 1e3, 3)
            out[f"{name}_count"] = int(sel.size)
    pix = sum(r.width * r.height for r in requests if r.kind == HD_REQ_TRACE)
    out["trace_mrays_per_s"] = round(pix / span / 1e6, 2)
    per_batch = {int(results["batch"][i]): int(results["voxels"][i])
                 for i in range(len(requests)) if requests[i].kind == HD_REQ_EDIT}
    out["edited_voxels"] = sum(per_batch.values())
    out["rejected"] = int((results["status"] != 0).sum())
    return out
